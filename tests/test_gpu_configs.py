"""GPU: the batched and large benchmark workloads at their own sizes.

* The 16 uneven cone-flat instances of the batched benchmark (volumes 128^2 x 32 to
  384^2 x 96 at 1 mm, detectors 2W x D, 180..984 views with random start angles; the
  instance list of ``bench.batched_instances``) run through ``BatchedProjector3d``: every
  operator in one launch per pass, per instance
    - bitwise equal to the plain per-instance execution at the same tile shape;
    - volume-in operators (A, B^T, B_W^T): four sampled views against the FP64 oracle
      run on those views;
    - sinogram-in operators (B, B_W, A^T): a sinogram that is zero outside four views
      against the FP64 oracle run on those four views (all three are sums over views).
  relative L2 <= 1e-5.
* Config 5 (cone-arc 512^2 x 256 at 0.7 mm, 736 x 128 arc detector, 46.6 degree fan) at
  full size: A^T, B^T, B_W^T on four views against the FP64 oracle; batch 4 over all 1152
  views, each batch item bitwise equal to its own single-item run (A and B); two
  checkpointed unrolled gradient iterations at batch 4 with finite gradients equal to the
  non-checkpointed ones.
"""
import numpy as np
import pytest
import torch

import oracle
import xrecon_amd
from helpers import OP_ID, rel_l2, uniform_angles
from xrecon_amd.batched import BatchedProjector3d

pytestmark = pytest.mark.gpu

CELL_OPS = ("A", "BT", "BWT")
VOXEL_OPS = ("B", "BW", "AT")


def _views(V):
    """Four views spread over the scan: the first, one near 45 degrees past the start
    (driving-axis switch region), one past half, the last."""
    return np.array([0, V // 8, V // 2 + 3, V - 1])


@pytest.fixture(scope="module")
def cfg3(cuda):
    import bench

    inst = bench.batched_instances()
    projs, angs = [], []
    for it in inst:
        W, D = it["W"], it["D"]
        projs.append(xrecon_amd.Projector3d((W, W, D), (it["Nu"], it["Nv"]), 1.0, 1.0, 500.0, 1000.0))
        angs.append(it["start"] + torch.arange(it["V"], dtype=torch.float32) * np.float32(2 * np.pi / it["V"]))
    bp = BatchedProjector3d(projs, angs, device=cuda)
    return projs, angs, bp


def _plain(bp, op_id, inputs, tiles, cuda):
    s = torch.cuda.current_stream(cuda).cuda_stream
    outs = []
    for i, plan in enumerate(bp.plans):
        out = torch.empty(bp.output_shape(op_id, i), device=cuda)
        ws = torch.empty(max(plan.workspace_bytes(op_id, 1), 1), dtype=torch.uint8, device=cuda)
        plan.execute_into(op_id, inputs[i], out, ws, s, tiles[0], tiles[1], 0)
        outs.append(out)
    return outs


@pytest.mark.parametrize("op", CELL_OPS)
def test_cfg3_batched_cell_ops(cfg3, cuda, op):
    projs, angs, bp = cfg3
    op_id = OP_ID[op]
    gen = torch.Generator(device=cuda).manual_seed(3100 + op_id)
    vols = [torch.rand(bp.volume_shape(i), device=cuda, generator=gen) for i in range(len(projs))]
    tiles = (16, 0)
    print(f"[gpu] cfg3 batched {op}: 16 instances", flush=True)
    outs = bp.run(op_id, vols, tiles=tiles)
    plain = _plain(bp, op_id, vols, tiles, cuda)
    torch.cuda.synchronize()
    for i, p in enumerate(projs):
        assert torch.equal(outs[i], plain[i]), f"instance {i}: batched differs from the plain launch"
        sel = _views(angs[i].numel())
        want = oracle.op3d(op, vols[i].cpu().numpy()[:, 0].astype(np.float64), angs[i].numpy()[sel],
                           p._host_options)
        got = outs[i].cpu().numpy()[:, 0][:, sel]
        err = rel_l2(got, want)
        assert err <= 1e-5, f"instance {i} ({p.image_size}, {angs[i].numel()} views): rel L2 {err:.2e}"


@pytest.mark.parametrize("op", VOXEL_OPS)
def test_cfg3_batched_voxel_ops(cfg3, cuda, op):
    projs, angs, bp = cfg3
    op_id = OP_ID[op]
    gen = torch.Generator(device=cuda).manual_seed(3200 + op_id)
    sinos, sels = [], []
    for i in range(len(projs)):
        sel = _views(angs[i].numel())
        y = torch.zeros(bp.sino_shape(i), device=cuda)
        y[:, :, torch.from_numpy(sel).to(cuda)] = torch.rand(y[:, :, :4].shape, device=cuda, generator=gen)
        sinos.append(y)
        sels.append(sel)
    tiles = (0, 16)
    print(f"[gpu] cfg3 batched {op}: 16 instances", flush=True)
    outs = bp.run(op_id, sinos, tiles=tiles)
    plain = _plain(bp, op_id, sinos, tiles, cuda)
    torch.cuda.synchronize()
    for i, p in enumerate(projs):
        assert torch.equal(outs[i], plain[i]), f"instance {i}: batched differs from the plain launch"
        y4 = sinos[i].cpu().numpy()[:, 0][:, sels[i]].astype(np.float64)
        want = oracle.op3d(op, y4, angs[i].numpy()[sels[i]], p._host_options)
        err = rel_l2(outs[i].cpu().numpy()[:, 0], want)
        assert err <= 1e-5, f"instance {i} ({p.image_size}, {angs[i].numel()} views): rel L2 {err:.2e}"


# ------------------------------------------------------------------------ config 5

CFG5 = [512, 512, 256, 736, 128, 0.7, 0.7, 0.7, np.float32(1.2 / 1085.6), 1.0, 595.0, 1085.6, 0, 0, 0, 0, 0, 2]


def _cfg5_projector():
    return xrecon_amd.Projector3d((512, 512, 256), (736, 128), 0.7, (1.2, 1.0), 595.0, 1085.6, scan_type="arc")


def test_cfg5_projector_options_match():
    o = _cfg5_projector()._host_options
    assert np.allclose(np.asarray(o, dtype=np.float64), np.asarray(CFG5, dtype=np.float32).astype(np.float64))


@pytest.mark.parametrize("op", ["AT", "BT", "BWT"])
def test_cfg5_arc_transposes_full_size(cuda, op):
    from xrecon_amd import _backend as be

    C = be.native()
    ang = uniform_angles(1152)
    sub = np.ascontiguousarray(ang[np.array([0, 144, 289, 577])])
    plan = C.Plan([float(v) for v in CFG5], torch.from_numpy(sub), 0, False)
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(5100 + OP_ID[op])
    if op == "AT":
        x = rng.random((1, 4, 128, 736)).astype(np.float32)
    else:
        x = rng.random((1, 256, 512, 512), dtype=np.float32)
    print(f"[gpu] cfg5 arc {op} on 4 views", flush=True)
    got = plan.execute(OP_ID[op], torch.from_numpy(x).to(cuda).unsqueeze(1), s).cpu().numpy()[:, 0]
    want = oracle.op3d(op, x.astype(np.float64), sub, CFG5)
    err = rel_l2(got, want)
    print(f"[gpu] cfg5 arc {op}: rel L2 {err:.2e}", flush=True)
    assert err <= 1e-5


def test_cfg5_staged_cell_gather(cuda):
    """Full size, 1152 views: the default cell gather stages each workgroup's blocks in
    LDS here (a z-prefix table of 539 MB per item), with some lines over a slot read from
    the table.  Bitwise equal to the direct gather; four sampled views against the oracle."""
    from xrecon_amd import _backend as be

    C = be.native()
    ang = uniform_angles(1152)
    plan = C.Plan([float(v) for v in CFG5], torch.from_numpy(ang), 0, False)
    s = torch.cuda.current_stream().cuda_stream
    x = np.random.default_rng(5200).random((1, 256, 512, 512), dtype=np.float32)
    xd = torch.from_numpy(x).to(cuda).unsqueeze(1)
    ws = torch.empty(plan.workspace_bytes(OP_ID["A"], 1), dtype=torch.uint8, device=cuda)
    outs = []
    for rows in (0, 8 | 0x100):  # default (staged), XR_CELL_DIRECT
        print(f"[gpu] cfg5 arc A, 1152 views, tile_rows {rows:#x}", flush=True)
        out = torch.empty((1, 1, 1152, 128, 736), device=cuda)
        plan.execute_into(OP_ID["A"], xd, out, ws, s, rows, 0, 0)
        outs.append(out)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    views = np.array([0, 144, 289, 577])
    want = oracle.op3d("A", x.astype(np.float64), np.ascontiguousarray(ang[views]), CFG5)
    err = rel_l2(outs[0][:, 0, views].cpu().numpy(), want)
    print(f"[gpu] cfg5 arc A staged, sampled views: rel L2 {err:.2e}", flush=True)
    assert err <= 1e-5


def test_cfg5_batch4_items_equal_single_runs(cuda):
    from xrecon_amd import _backend as be

    C = be.native()
    ang = torch.from_numpy(uniform_angles(1152))
    plan = C.Plan([float(v) for v in CFG5], ang, 0, False)
    s = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device=cuda).manual_seed(55)
    x = torch.rand((4, 1, 256, 512, 512), device=cuda, generator=g)
    print("[gpu] cfg5 arc A, batch 4, 1152 views", flush=True)
    y4 = plan.execute(OP_ID["A"], x, s)
    for b in range(4):
        y1 = plan.execute(OP_ID["A"], x[b:b + 1].contiguous(), s)
        assert torch.equal(y4[b], y1[0]), f"A item {b}"
        del y1
    del x
    print("[gpu] cfg5 arc B, batch 4, 1152 views", flush=True)
    r4 = plan.execute(OP_ID["B"], y4, s)
    for b in range(4):
        r1 = plan.execute(OP_ID["B"], y4[b:b + 1].contiguous(), s)
        assert torch.equal(r4[b], r1[0]), f"B item {b}"
    assert torch.isfinite(r4).all()


def test_cfg5_checkpointed_unroll_two_iterations(cuda):
    """Two unrolled gradient iterations x <- x - a_k A^T (A x - y) at config 5 size,
    batch 4, 1152 views: gradients of the step sizes are finite and the same with and
    without per-iteration checkpointing."""
    from xrecon_amd.recon import UnrolledGradientDescent

    proj = _cfg5_projector()
    angles = torch.from_numpy(uniform_angles(1152)).to(cuda)
    g = torch.Generator(device=cuda).manual_seed(56)
    xt = 0.05 * torch.rand((4, 1, 256, 512, 512), device=cuda, generator=g)
    with torch.no_grad():
        y = proj.projection(xt, angles)
    grads, outs = [], []
    for ckpt in (True, False):
        print(f"[gpu] cfg5 unrolled, 2 iterations, checkpoint={ckpt}", flush=True)
        net = UnrolledGradientDescent(proj, angles, iterations=2, step=1e-6, use_checkpoint=ckpt).to(cuda)
        out = net(y)
        loss = ((out - xt) ** 2).sum()
        loss.backward()
        grads.append(net.steps.grad.clone())
        outs.append(out.detach())
        del out, loss, net
        torch.cuda.synchronize()
    assert torch.isfinite(grads[0]).all()
    assert torch.equal(outs[0], outs[1])
    assert torch.equal(grads[0], grads[1])
