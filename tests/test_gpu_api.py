"""GPU: the public API end to end at the benchmark configurations.

* FDK (Projector3d.filtered_backprojection) and 2D FBP (Projector2d) against an FP64
  pipeline built from the projector's own weight and filter buffers, a direct FP64
  correlation (the reference's zero-padded conv, projector3d.py:190 / projector.py:201)
  and the FP64 oracle's weighted backprojection: relative L2 <= 1e-5.
* The fused FDK pass (xr_plan3d_execute_fdk: filter stage, in-place prefix, weighted
  voxel gather) on every geometry against the FP64 pipeline and the CPU emulator of the
  same kernel bodies, and the projector's fused and differentiable composite paths
  against each other.
* Config 1 (2D fan-flat 256^2, 256 detectors, 720 views) and config 5 (cone-arc
  512^2 x 256 at 0.7 mm, 736 x 128 detector, 46.6 degree fan) at full size: A on sampled
  views and B, B_W on a subset of views against the FP64 oracle, relative L2 <= 1e-5.
* Config 4: the transmission simulation round trip and a few reconstruction steps.
"""
import numpy as np
import pytest
import torch

import oracle
import xrecon_amd
from helpers import CASES3D, angles_with_ties, correlate_fp64, emu_available, emu_fdk3d, rel_l2, shapes3d, \
    uniform_angles

pytestmark = pytest.mark.gpu


def _mean_step(ang):
    d = np.diff(ang.astype(np.float64))
    return np.abs((d + np.pi) % (2 * np.pi) - np.pi).mean()


def test_fdk_pipeline_vs_fp64(cuda):
    proj = xrecon_amd.Projector3d((48, 44, 12), (96, 16), 1.0, 1.0, 500.0, 1000.0, scan_type="flat").to(cuda)
    V = 90
    ang = uniform_angles(V)
    rng = np.random.default_rng(21)
    y = rng.random((1, 1, V, 16, 96)).astype(np.float32)
    got = proj.filtered_backprojection(torch.from_numpy(y).to(cuda), torch.from_numpy(ang).to(cuda))
    w = proj.weight.detach().cpu().double().numpy().reshape(-1)
    h = proj.filter.detach().cpu().double().numpy().reshape(-1)
    filtered = correlate_fp64(y[0, 0].astype(np.float64) * w, h)
    want = oracle.op3d("BW", filtered[None], ang, list(proj._host_options)) * _mean_step(ang) / 2
    assert rel_l2(got.cpu().numpy()[0, 0], want[0]) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES3D))
def test_fused_fdk_vs_fp64_and_emulator(cuda, geo):
    from xrecon_amd import _backend as be
    from xrecon_amd.filters import filter_generate

    opts = CASES3D[geo]
    V = 40
    ang = angles_with_ties(V)
    _, sino = shapes3d(opts, 2, V)
    nu = sino[-1]
    rng = np.random.default_rng(41)
    filt = (filter_generate(nu, 1.0, "shepplogan").numpy() * 0.37 + 0.05 * rng.random(2 * nu - 1)).astype(np.float32)
    weight = (0.5 + rng.random(nu)).astype(np.float32)
    y = rng.random(sino).astype(np.float32)
    plan = be.native().Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    print(f"[gpu] fused FDK {geo} shape={y.shape}", flush=True)
    got = plan.fdk(torch.from_numpy(y).to(cuda).unsqueeze(1), torch.from_numpy(filt).to(cuda),
                   torch.from_numpy(weight).to(cuda), torch.cuda.current_stream().cuda_stream)
    got = got.cpu().numpy()[:, 0]
    want = oracle.op3d("BW", correlate_fp64(y.astype(np.float64) * weight, filt), ang, opts)
    assert rel_l2(got, want) <= 1e-5
    if emu_available():
        assert rel_l2(got, emu_fdk3d(y, ang, opts, filt, weight)) <= 1e-6


@pytest.mark.parametrize("scan", ["flat", "arc"])
def test_fused_fdk_matches_differentiable_path(cuda, scan):
    """filtered_backprojection without gradients (fused pass) and with a gradient wanted
    through the sinogram (weight, FFT correlation, weighted backprojection)."""
    proj = xrecon_amd.Projector3d((40, 36, 10), (64, 12), 1.0, 1.0, 500.0, 1000.0, scan_type=scan).to(cuda)
    ang = torch.from_numpy(uniform_angles(72)).to(cuda)
    y = torch.rand((2, 1, 72, 12, 64), device=cuda, generator=torch.Generator(device=cuda).manual_seed(9))
    with torch.no_grad():
        fused = proj.filtered_backprojection(y, ang)
    yg = y.clone().requires_grad_(True)
    comp = proj.filtered_backprojection(yg, ang)
    assert comp.requires_grad
    assert rel_l2(fused.cpu().numpy(), comp.detach().cpu().numpy()) <= 1e-5


def test_fbp2d_config1_vs_fp64(cuda):
    """Config 1 at full size: A of a phantom and the FBP of it, both against FP64."""
    from xrecon_amd.phantom import random_ellipses

    proj = xrecon_amd.Projector2d(256, 256, 1.0, 2.0, 500.0, 1000.0, scan_type="flat").to(cuda)
    V = 720
    ang = uniform_angles(V)
    img = random_ellipses((256, 256), (1.0, 1.0), count=10, semi_axes=(5.0, 60.0), seed=1001)
    opts = list(proj._host_options)
    a = torch.from_numpy(ang).to(cuda)
    sino = proj.projection(torch.from_numpy(img).to(cuda).view(1, 1, 256, 256), a)
    want_sino = oracle.op2d("A", img[None].astype(np.float64), ang, opts)
    assert rel_l2(sino.cpu().numpy()[0, 0], want_sino[0]) <= 1e-5
    rec = proj.filtered_backprojection(sino, a)
    w = proj.weight.detach().cpu().double().numpy().reshape(-1)
    h = proj.filter.detach().cpu().double().numpy().reshape(-1)
    filtered = correlate_fp64(sino.cpu().double().numpy()[0, 0] * w, h)
    want = oracle.op2d("BW", filtered[None], ang, opts) * _mean_step(ang) / 2
    assert rel_l2(rec.cpu().numpy()[0, 0], want[0]) <= 1e-5


@pytest.mark.parametrize("scan", ["arc", "parallel"])
def test_2d_config1_size_fan_arc_and_parallel(cuda, scan):
    """Config 1's image and sampling (256^2 at 1 mm, 256 detectors, 720 views) on the other
    2D geometries: A and B_W at full size against the FP64 oracle."""
    from xrecon_amd.phantom import random_ellipses

    if scan == "arc":  # the 2 mm detector on an arc of radius SDD: 2 mm / 1000 mm per column
        proj = xrecon_amd.Projector2d(256, 256, 1.0, 2.0, 500.0, 1000.0, scan_type="arc").to(cuda)
    else:
        proj = xrecon_amd.Projector2d(256, 256, 1.0, 1.0, 0.0, 0.0, scan_type="parallel").to(cuda)
    V = 720
    ang = uniform_angles(V)
    opts = list(proj._host_options)
    img = random_ellipses((256, 256), (1.0, 1.0), count=10, semi_axes=(5.0, 60.0), seed=1001)
    a = torch.from_numpy(ang).to(cuda)
    print(f"[gpu] 2D {scan} config-1 size A, B_W", flush=True)
    sino = proj.projection(torch.from_numpy(img).to(cuda).view(1, 1, 256, 256), a)
    want = oracle.op2d("A", img[None].astype(np.float64), ang, opts)
    assert rel_l2(sino.cpu().numpy()[0, 0], want[0]) <= 1e-5
    y = np.random.default_rng(31).random((1, V, 256)).astype(np.float32)
    got = proj.backprojection_weighted(torch.from_numpy(y).to(cuda).unsqueeze(1), a)
    want = oracle.op2d("BW", y.astype(np.float64), ang, opts)
    assert rel_l2(got.cpu().numpy()[0, 0], want[0]) <= 1e-5


CFG5 = [512, 512, 256, 736, 128, 0.7, 0.7, 0.7, np.float32(1.2 / 1085.6), 1.0, 595.0, 1085.6, 0, 0, 0, 0, 0, 2]


def test_cfg5_arc_full_size_vs_oracle(cuda):
    """Config 5 geometry at full size (one batch item): A on six views including the
    45-degree driving-axis ties, B and B_W on four views, against the FP64 oracle."""
    from xrecon_amd import _backend as be
    from xrecon_amd.phantom import random_ellipsoids

    C = be.native()
    V = 1152
    ang = uniform_angles(V)
    views = np.array([0, 1, 144, 288, 577, 1000])
    sub = np.ascontiguousarray(ang[views])
    ph = random_ellipsoids((512, 512, 256), (0.7, 0.7, 0.7), count=16, semi_axes=(5.0, 120.0), seed=1005)
    plan = C.Plan([float(v) for v in CFG5], torch.from_numpy(sub), 0, False)
    s = torch.cuda.current_stream().cuda_stream
    print("[gpu] cfg5 arc A on 6 views", flush=True)
    y = plan.execute(0, torch.from_numpy(ph).to(cuda).view(1, 1, 256, 512, 512), s)
    want = oracle.op3d("A", ph[None].astype(np.float64), sub, CFG5)
    assert rel_l2(y.cpu().numpy()[0, 0], want[0]) <= 1e-5
    del ph
    bviews = np.array([0, 144, 288, 577])
    bsub = np.ascontiguousarray(ang[bviews])
    pb = C.Plan([float(v) for v in CFG5], torch.from_numpy(bsub), 0, False)
    yb = np.random.default_rng(23).random((1, 4, 128, 736)).astype(np.float32)
    for op, opid in (("B", 2), ("BW", 3)):
        print(f"[gpu] cfg5 arc {op} on 4 views", flush=True)
        got = pb.execute(opid, torch.from_numpy(yb).to(cuda).unsqueeze(1), s).cpu().numpy()[:, 0]
        want = oracle.op3d(op, yb.astype(np.float64), bsub, CFG5)
        assert rel_l2(got, want) <= 1e-5, op


def test_cfg4_transmission_and_reconstruction_steps(cuda):
    """Config 4 in miniature: y = -ln(max(Poisson(I0 e^{-A x}), 1) / I0) is finite and
    close to A x at high dose; gradient steps of size 1/L (L from a power iteration on
    A^T A) lower 0.5 ||A z - y||^2 monotonically; the Adam reconstruction runs."""
    from xrecon_amd.recon import data_loss, reconstruct, simulate_transmission

    proj = xrecon_amd.Projector3d((32, 32, 8), (48, 8), 1.0, 1.0, 500.0, 1000.0)
    ang = torch.from_numpy(uniform_angles(60)).to(cuda)
    x = torch.zeros((1, 1, 8, 32, 32), device=cuda)
    x[..., 2:6, 8:24, 8:24] = 0.02
    y = simulate_transmission(proj, x, ang, I0=1e7, seed=3)
    clean = proj.projection(x, ang)
    assert torch.isfinite(y).all()
    assert rel_l2(y.cpu().numpy(), clean.cpu().numpy()) <= 2e-2
    with torch.no_grad():
        v = torch.rand_like(x)
        for _ in range(10):
            w = proj.projection_t(proj.projection(v, ang), ang)
            lip = (w.norm() / v.norm()).item()
            v = w / w.norm()
        tau = 1.0 / (1.1 * lip)
        z = torch.zeros_like(x)
        losses = [data_loss(proj, z, y, ang).item()]
        for _ in range(5):
            z = z - tau * proj.projection_t(proj.projection(z, ang) - y, ang)
            losses.append(data_loss(proj, z, y, ang).item())
    assert all(b <= a for a, b in zip(losses, losses[1:])), losses
    assert losses[-1] < 0.9 * losses[0], losses
    z = reconstruct(proj, y, ang, steps=3, lr=1e-4, tv_weight=1e-3)
    assert torch.isfinite(z).all()
