"""Seeded random geometries: every operator against the FP64 oracle (relative L2 <= 1e-5),
on the CPU emulator of the kernel bodies and on the GPU, with the default tiles and with
the halved voxel chunks.  Sizes, spacings, shifts, detector pitch and distances, view
counts and angle sets (uniform with a random start, or sorted random angles) are drawn
per seed, so the kernels' tile edges, span classes, staging blocks, view sets and slice
chunks meet shapes no hand-written case has.  On the emulator the cell gathers also run
as LDS-staged workgroups, bit for bit equal to the direct threads."""
import numpy as np
import pytest
import torch

import oracle
from helpers import OP_ID, OPS, VOLUME_IN, emu_available, emu_cell_lines, emu_op3d, opts3d, rel_l2, shapes3d

SEEDS = list(range(12))


def random_case(seed):
    rng = np.random.default_rng(9000 + seed)
    scan = ("flat", "arc", "par")[seed % 3]
    W, H = (int(v) for v in rng.integers(8, 71, 2))
    D = int(rng.integers(1, 41))
    dx, dy, dz = (float(v) for v in rng.uniform(0.5, 1.5, 3))
    Nu, Nv = int(rng.integers(16, 121)), int(rng.integers(1, 41))
    dv = float(rng.uniform(0.5, 1.5))
    if scan == "par":
        sod = sdd = 0.0
        du = float(rng.uniform(0.5, 1.5))
        shu = float(rng.uniform(-0.5, 0.5)) * du
    else:
        sod = float(rng.uniform(300.0, 700.0))
        sdd = sod * float(rng.uniform(1.5, 2.5))
        du = float(rng.uniform(0.6, 1.6))
        shu = float(rng.uniform(-0.5, 0.5)) * du
        if scan == "arc":  # the kernel convention: radians (du / SDD)
            du, shu = du / sdd, shu / sdd
    shx, shy, shz = (float(v) for v in rng.uniform(-0.5, 0.5, 3) * np.array([dx, dy, dz]))
    shv = float(rng.uniform(-0.5, 0.5)) * dv
    opts = opts3d(W, H, D, Nu, Nv, dx, dy, dz, du, dv, sod, sdd, {"par": 0, "flat": 1, "arc": 2}[scan],
                  shx, shy, shz, shu, shv)
    opts = [float(np.float32(v)) for v in opts]
    V = int(rng.integers(5, 41))
    if seed % 2:
        ang = np.float32(rng.uniform(0, 2 * np.pi)) + np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V)
    else:
        ang = np.sort(rng.uniform(0, 2 * np.pi, V))
    return opts, np.asarray(ang, dtype=np.float32), rng


@pytest.mark.skipif(not emu_available(), reason="emulator not built (run __graft_entry__.build())")
@pytest.mark.parametrize("seed", SEEDS)
def test_random_geometry_bodies_vs_oracle(seed):
    opts, ang, rng = random_case(seed)
    vol, sino = shapes3d(opts, 1, ang.shape[0])
    for op in OPS:
        x = rng.random(vol if op in VOLUME_IN else sino).astype(np.float32)
        want = oracle.op3d(op, x.astype(np.float64), ang, opts)
        got = emu_op3d(op, x, ang, opts)
        assert rel_l2(got, want) <= 1e-5, (seed, op)
        if op in VOLUME_IN:  # the cell gathers' LDS-staged workgroups too, bit for bit
            for R in (4, 16):
                emu_cell_lines()
                staged = emu_op3d(op, x, ang, opts, R=R, cell_variant=3)
                assert emu_cell_lines()[1] == 0, (seed, op, R)
                assert np.array_equal(staged, emu_op3d(op, x, ang, opts, R=R)), (seed, op, R)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
def test_random_geometry_gpu_vs_oracle(cuda, seed):
    from xrecon_amd import _backend as be

    opts, ang, rng = random_case(seed)
    plan = be.native().Plan(opts, torch.from_numpy(ang), 0, False)
    s = torch.cuda.current_stream().cuda_stream
    vol, sino = shapes3d(opts, 2, ang.shape[0])
    for op in OPS:
        x = rng.random(vol if op in VOLUME_IN else sino).astype(np.float32)
        want = oracle.op3d(op, x.astype(np.float64), ang, opts)
        xd = torch.from_numpy(x).to(cuda).unsqueeze(1)
        k = OP_ID[op]
        out = plan.execute(k, xd, s)
        ws = torch.empty(max(plan.workspace_bytes(k, 2), 1), dtype=torch.uint8, device=cuda)
        tiles = [(0, 0)] if op in VOLUME_IN else [(0, 0), (0, 16 | 0x100), (0, 32 | 0x100)]
        for rows, chunk in tiles:
            print(f"[gpu] random geometry {seed} {op} chunk {chunk:#x}", flush=True)
            plan.execute_into(k, xd, out, ws, s, rows, chunk, 0)
            assert rel_l2(out.cpu().numpy()[:, 0], want) <= 1e-5, (seed, op, chunk)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
def test_random_geometry_gpu_staged_equals_direct(cuda, seed):
    """The LDS-staged cell gather on the GPU against the direct one, bit for bit: a batch
    large enough that the two-pixel class launches at least 4096 workgroups (staging
    applies to such launches only), 8-row tiles."""
    from xrecon_amd import _backend as be

    opts, ang, rng = random_case(seed)
    plan = be.native().Plan(opts, torch.from_numpy(ang), 0, False)
    n2 = plan.cell_classes()[0]
    nv = int(opts[4])
    if n2 == 0:
        pytest.skip("no two-pixel column groups in this geometry")
    per_item = -(-n2 // 4) * -(-nv // 8)
    nb = min(1024, -(-4096 // per_item))
    vol, _ = shapes3d(opts, nb, ang.shape[0])
    x = torch.from_numpy(rng.random(vol).astype(np.float32)).to(cuda).unsqueeze(1)
    s = torch.cuda.current_stream().cuda_stream
    for op in VOLUME_IN:
        k = OP_ID[op]
        ws = torch.empty(max(plan.workspace_bytes(k, nb), 1), dtype=torch.uint8, device=cuda)
        print(f"[gpu] random geometry {seed} {op} staged vs direct, batch {nb}", flush=True)
        outs = []
        for rows in (8 | 0x200, 8 | 0x100):  # XR_CELL_LDS, XR_CELL_DIRECT
            out = plan.execute(k, x, s)
            plan.execute_into(k, x, out, ws, s, rows, 0, 0)
            outs.append(out)
        torch.cuda.synchronize()
        assert torch.equal(outs[0], outs[1]), (seed, op)
