"""GPU: batched single-launch execution over uneven instances, and graph capture.

A batched launch runs the same per-thread bodies as the single-instance kernels, so with
the same tile shape its outputs must be bitwise identical to per-instance execution.
Graph replays are checked against the eager result on every replay, with eager work and
synchronisations between replays."""
import numpy as np
import pytest
import torch

import oracle
import xrecon_amd
from helpers import rel_l2
from xrecon_amd import _backend as be
from xrecon_amd.batched import BatchedProjector3d

pytestmark = pytest.mark.gpu

SHAPES = [((24, 20, 8), (40, 8), 17, 0.3, 1), ((32, 32, 12), (64, 12), 23, 1.1, 2),
          ((16, 28, 6), (32, 6), 11, 2.5, 1), ((40, 36, 10), (72, 10), 29, 4.0, 1)]


def _instances(scan="flat"):
    projs, angs = [], []
    for (img, det, V, start, _) in SHAPES:
        du = 1.0 if scan == "flat" else 1.0 / 1000.0
        projs.append(xrecon_amd.Projector3d(img, det, 1.0, (du, 1.0), 500.0, 1000.0, scan_type=scan))
        angs.append((np.float32(start) + torch.arange(V, dtype=torch.float32) * np.float32(2 * np.pi / V)).float())
    return projs, angs, [s[4] for s in SHAPES]


# (cell rows, voxel chunk): plain tiles, and 16 / 32 slices in wave-culled halves
TILES = [(4, 4), (16, 16 | 0x100), (8, 32 | 0x100)]


@pytest.mark.parametrize("scan", ["flat", "arc"])
@pytest.mark.parametrize("op", [be.OP_PROJECTION, be.OP_BACKPROJECTION, be.OP_PROJECTION_T,
                                be.OP_BACKPROJECTION_W, be.OP_BACKPROJECTION_T])
@pytest.mark.parametrize("tile", range(len(TILES)))
def test_batched_equals_per_instance(cuda, scan, op, tile):
    projs, angs, nbs = _instances(scan)
    bp = BatchedProjector3d(projs, angs, nbs, device=cuda)
    gen = torch.Generator(device=cuda).manual_seed(op)
    cell = bp.plans[0].kernel_kind(op) == 1
    inputs = [torch.rand(bp.volume_shape(i) if cell else bp.sino_shape(i), device=cuda, generator=gen)
              for i in range(len(projs))]
    tiles = (TILES[tile][0], 0) if cell else (0, TILES[tile][1])
    print(f"[gpu] batch {scan} op {op} tiles {tiles}", flush=True)
    outs = bp.run(op, inputs, tiles=tiles)
    s = torch.cuda.current_stream().cuda_stream
    for i, plan in enumerate(bp.plans):
        single = torch.empty_like(outs[i])
        ws = torch.empty(max(plan.workspace_bytes(op, nbs[i]), 1), dtype=torch.uint8, device=cuda)
        plan.execute_into(op, inputs[i], single, ws, s, tiles[0], tiles[1], 0)
        assert torch.equal(outs[i], single), f"instance {i}"


def test_batched_vs_oracle(cuda):
    projs, angs, nbs = _instances("flat")
    bp = BatchedProjector3d(projs, angs, nbs, device=cuda)
    rng = np.random.default_rng(0)
    vols = [rng.random(bp.volume_shape(i)).astype(np.float32) for i in range(len(projs))]
    outs = bp.run(be.OP_PROJECTION, [torch.from_numpy(v).to(cuda) for v in vols])
    for i, p in enumerate(projs):
        want = oracle.op3d("A", vols[i][:, 0].astype(np.float64), angs[i].numpy(), p._host_options)
        assert rel_l2(outs[i].cpu().numpy()[:, 0], want) <= 1e-5


def test_graph_capture_replays_match_eager(cuda):
    projs, angs, nbs = _instances("flat")
    bp = BatchedProjector3d(projs, angs, nbs, device=cuda)
    gen = torch.Generator(device=cuda).manual_seed(1)
    vols = [torch.rand(bp.volume_shape(i), device=cuda, generator=gen) for i in range(len(projs))]
    ops = [be.OP_PROJECTION, be.OP_BACKPROJECTION]
    eager_a = bp.run(ops[0], vols)
    eager_b = bp.run(ops[1], eager_a)
    bufs = bp.capture(ops, vols)
    for rep in range(3):
        bp.replay()
        torch.cuda.synchronize()
        for i in range(len(projs)):
            assert torch.equal(bufs[0][i], eager_a[i]), f"replay {rep} A instance {i}"
            assert torch.equal(bufs[1][i], eager_b[i]), f"replay {rep} B instance {i}"
        # eager work between replays
        junk = bp.run(ops[0], vols)
        torch.cuda.synchronize()
        del junk
    # new data in the static input buffers is picked up by the next replay
    for v in vols:
        v.mul_(2.0)
    bp.replay()
    torch.cuda.synchronize()
    for i in range(len(projs)):
        assert torch.allclose(bufs[0][i], 2.0 * eager_a[i], rtol=1e-6, atol=0)


def test_batch_rejects_mixed_geometry(cuda):
    p1 = xrecon_amd.Projector3d((8, 8, 4), (16, 4), 1.0, 1.0, 500.0, 1000.0, scan_type="flat")
    p2 = xrecon_amd.Projector3d((8, 8, 4), (16, 4), 1.0, (0.001, 1.0), 500.0, 1000.0, scan_type="arc")
    a = torch.arange(4, dtype=torch.float32)
    with pytest.raises(RuntimeError, match="geometry family"):
        BatchedProjector3d([p1, p2], [a, a], device=cuda)
