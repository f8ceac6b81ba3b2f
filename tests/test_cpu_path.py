"""CPU: the host executor for CPU tensors (xr_host_plan3d_*, the counterpart of the
reference's CPU dispatch, src/api/torch_binding.cpp:421-425).

It runs the product's own per-thread bodies on the host, so it is checked like the GPU
path: every operator and geometry against the FP64 oracle (relative L2 <= 1e-5), through
the reference-named ``_C`` functions, the projector API (including FDK/FBP), autograd
(the gradient is the adjoint, dot-product test), the one-shot C-ABI, and against the
reference's own CPU kernels where oracle/_ref is built."""
import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
import xrecon_amd
from helpers import CASES2D, CASES3D, OPS, VOLUME_IN, angles_with_ties, correlate_fp64, rel_l2, shapes3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "xrecon_amd", "_lib", "libxrecon_amd.so")

NAMES3D = {"A": "forward_projection3d", "AT": "projection_transpose3d", "B": "backward_projection3d",
           "BW": "weighted_backward_projection3d", "BT": "backward_projection_transpose3d",
           "BWT": "weighted_backward_projection_transpose3d"}
NAMES2D = {"forward_projection2d": "A", "projection_transpose2d": "AT", "backward_projection2d": "B",
           "weighted_backward_projection2d": "BW", "backward_projection_transpose2d": "BT",
           "weighted_backward_projection_transpose2d": "BWT", "back_projection2d": "B",
           "weighted_back_projection2d": "BW"}


def _C():
    from xrecon_amd import _backend as be

    return be.native()


@pytest.mark.parametrize("geo", sorted(CASES3D))
@pytest.mark.parametrize("op", OPS)
def test_reference_named_3d_on_cpu_tensors(geo, op):
    opts = CASES3D[geo]
    V = 24
    ang = angles_with_ties(V)
    vol, sino = shapes3d(opts, 2, V)
    x = np.random.default_rng(500 + OPS.index(op)).random(vol if op in VOLUME_IN else sino).astype(np.float32)
    got = getattr(_C(), NAMES3D[op])(torch.from_numpy(x).unsqueeze(1), torch.tensor(opts, dtype=torch.float32),
                                     torch.from_numpy(ang))
    assert not got.is_cuda
    assert rel_l2(got.numpy()[:, 0], oracle.op3d(op, x.astype(np.float64), ang, opts)) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES2D))
@pytest.mark.parametrize("name", sorted(NAMES2D))
def test_reference_named_2d_on_cpu_tensors(geo, name):
    op = NAMES2D[name]
    o2 = CASES2D[geo]
    W, H, Nu = int(o2[0]), int(o2[1]), int(o2[2])
    V = 30
    ang = angles_with_ties(V)
    x = np.random.default_rng(600 + OPS.index(op)).random((2, H, W) if op in VOLUME_IN else (2, V, Nu)).astype(np.float32)
    got = getattr(_C(), name)(torch.from_numpy(x).unsqueeze(1), torch.tensor(o2, dtype=torch.float32),
                            torch.from_numpy(ang))
    assert rel_l2(got.numpy()[:, 0], oracle.op2d(op, x.astype(np.float64), ang, o2)) <= 1e-5


@pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU library not built")
@pytest.mark.parametrize("geo", ["flat", "arc"])
def test_against_the_reference_cpu_kernels(geo):
    opts = CASES3D[geo]
    ang = angles_with_ties(20)
    vol, sino = shapes3d(opts, 1, 20)
    rng = np.random.default_rng(7)
    x, y = rng.random(vol).astype(np.float32), rng.random(sino).astype(np.float32)
    C = _C()
    o = torch.tensor(opts, dtype=torch.float32)
    a = torch.from_numpy(ang)
    ours_a = C.forward_projection3d(torch.from_numpy(x).unsqueeze(1), o, a).numpy()[:, 0]
    ours_b = C.backward_projection3d(torch.from_numpy(y).unsqueeze(1), o, a).numpy()[:, 0]
    tol = 2e-5 if geo == "flat" else 2e-4  # the reference's own FP32 tanf(beta + gamma) drift on arcs
    assert rel_l2(ours_a, oracle.ref_op3d("A", x, ang, opts)) <= tol
    assert rel_l2(ours_b, oracle.ref_op3d("B", y, ang, opts)) <= tol


def test_projector_api_fdk_and_autograd_on_cpu():
    proj = xrecon_amd.Projector3d((40, 36, 10), (64, 12), 1.0, 1.0, 500.0, 1000.0, scan_type="flat")
    V = 36
    ang = torch.arange(V, dtype=torch.float32) * np.float32(2 * np.pi / V)
    g = torch.Generator().manual_seed(3)
    x = torch.rand((1, 1, 10, 36, 40), generator=g)
    y = torch.rand((1, 1, V, 12, 64), generator=g)
    opts = list(proj._host_options)
    # projection and FDK against FP64
    sino = proj.projection(x, ang)
    assert rel_l2(sino.numpy()[0, 0], oracle.op3d("A", x.numpy()[:, 0].astype(np.float64), ang.numpy(), opts)[0]) <= 1e-5
    rec = proj.filtered_backprojection(y, ang)
    w = proj.weight.detach().double().numpy().reshape(-1)
    h = proj.filter.detach().double().numpy().reshape(-1)
    d = np.diff(ang.numpy().astype(np.float64))
    step = np.abs((d + np.pi) % (2 * np.pi) - np.pi).mean()
    want = oracle.op3d("BW", correlate_fp64(y.numpy()[0, 0].astype(np.float64) * w, h)[None], ang.numpy(), opts)
    assert rel_l2(rec.numpy()[0, 0], want[0] * step / 2) <= 1e-5
    # autograd: the gradient of 0.5 ||A x - y||^2 is A^T (A x - y), bitwise
    xg = x.clone().requires_grad_(True)
    loss = 0.5 * ((proj.projection(xg, ang) - y) ** 2).sum()
    (grad,) = torch.autograd.grad(loss, xg)
    with torch.no_grad():
        assert torch.equal(grad, proj.projection_t(proj.projection(x, ang) - y, ang))
    # dot-product test <A x, y> = <x, A^T y>
    lhs = (proj.projection(x, ang).double() * y.double()).sum().item()
    rhs = (x.double() * proj.projection_t(y, ang).double()).sum().item()
    assert abs(lhs - rhs) / abs(rhs) <= 1e-5


class _Vol3(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("dx", "dy", "dz", "shift_x", "shift_y", "shift_z")]


class _Geo3(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("nu", ctypes.c_int), ("nv", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("du", "dv", "shift_u", "shift_v", "sod", "sdd")]


class _Vol2(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("dx", "dy", "shift_x", "shift_y")]


class _Geo2(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("nu", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("du", "shift_u", "sod", "sdd")]


ONESHOT = {"A": ("xr_forward3d_cpu", True), "AT": ("xr_forward3d_t_cpu", True), "B": ("xr_backward3d_cpu", True),
           "BW": ("xr_weighted_backward3d_cpu", True), "BT": ("xr_backward3d_t_cpu", False),
           "BWT": ("xr_weighted_backward3d_t_cpu", False)}


@pytest.mark.parametrize("geo", sorted(CASES3D))
@pytest.mark.parametrize("op", OPS)
def test_oneshot_host_c_abi(geo, op):
    lib = ctypes.CDLL(LIB)
    lib.xr_last_error.restype = ctypes.c_char_p
    o = CASES3D[geo]
    V, nb = 14, 2
    ang = np.ascontiguousarray(angles_with_ties(V))
    vol, sino = shapes3d(o, nb, V)
    x = np.random.default_rng(700 + OPS.index(op)).random(vol if op in VOLUME_IN else sino).astype(np.float32)
    out = np.full(sino if op in VOLUME_IN else vol, np.nan, dtype=np.float32)
    v3 = _Vol3(int(o[0]), int(o[1]), int(o[2]), o[5], o[6], o[7], o[12], o[13], o[14])
    g3 = _Geo3(int(o[17]), int(o[3]), int(o[4]), o[8], o[9], o[15], o[16], o[10], o[11])
    name, image_first = ONESHOT[op]
    image, proj = (x, out) if op in VOLUME_IN else (out, x)
    first, second = (image, proj) if image_first else (proj, image)
    rc = getattr(lib, name)(first.ctypes.data_as(ctypes.c_void_p), second.ctypes.data_as(ctypes.c_void_p),
                            ang.ctypes.data_as(ctypes.c_void_p), ctypes.byref(v3), ctypes.byref(g3), nb, V)
    assert rc == 0, lib.xr_last_error()
    assert rel_l2(out, oracle.op3d(op, x.astype(np.float64), ang, o)) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES2D))
def test_oneshot_host_c_abi_2d(geo):
    lib = ctypes.CDLL(LIB)
    o2 = CASES2D[geo]
    W, H, Nu = int(o2[0]), int(o2[1]), int(o2[2])
    V = 20
    ang = np.ascontiguousarray(angles_with_ties(V))
    x = np.random.default_rng(8).random((1, H, W)).astype(np.float32)
    out = np.zeros((1, V, Nu), dtype=np.float32)
    v2 = _Vol2(W, H, o2[3], o2[4], o2[8], o2[9])
    g2 = _Geo2(int(o2[11]), Nu, o2[5], o2[10], o2[6], o2[7])
    rc = lib.xr_op2d_cpu(0, x.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                         ang.ctypes.data_as(ctypes.c_void_p), ctypes.byref(v2), ctypes.byref(g2), 1, V)
    assert rc == 0
    assert rel_l2(out, oracle.op2d("A", x.astype(np.float64), ang, o2)) <= 1e-5
