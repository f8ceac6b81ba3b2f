"""GPU: the drop-in boundary computes.

* Every reference-named function of ``xrecon_amd._C`` (the six 3D operators, the six 2D
  operators and the two 2D aliases) called the way the reference's
  ``xrecon/projector/functions.py:7-93`` calls ``xrecon_C``: ``(input, options, angles)``
  with the options vector and the angles as GPU tensors, against the FP64 oracle on every
  geometry (relative L2 <= 1e-5).
* Those functions resolve plans from a cache: a repeated call with the same tensors, or
  with equal geometry in new tensors, builds no new plan; by default they run on torch's
  current stream.
* The one-shot C-ABI entry points (``xr_forward3d`` ... ``xr_weighted_backward3d_t``,
  ``xr_op2d``; the reference's ``include/functions.h`` operator API) called through ctypes
  on device buffers, against the FP64 oracle.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

import oracle
from helpers import CASES2D, CASES3D, OPS, VOLUME_IN, angles_with_ties, rel_l2, shapes3d

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "xrecon_amd", "_lib", "libxrecon_amd.so")

NAMES3D = {"A": "forward_projection3d", "AT": "projection_transpose3d", "B": "backward_projection3d",
           "BW": "weighted_backward_projection3d", "BT": "backward_projection_transpose3d",
           "BWT": "weighted_backward_projection_transpose3d"}
NAMES2D = {"A": "forward_projection2d", "AT": "projection_transpose2d", "B": "backward_projection2d",
           "BW": "weighted_backward_projection2d", "BT": "backward_projection_transpose2d",
           "BWT": "weighted_backward_projection_transpose2d"}
ALIASES2D = {"back_projection2d": "B", "weighted_back_projection2d": "BW"}


def _C():
    from xrecon_amd import _backend as be

    return be.native()


@pytest.mark.parametrize("geo", sorted(CASES3D))
@pytest.mark.parametrize("op", OPS)
def test_reference_named_3d_functions_vs_oracle(cuda, geo, op):
    opts = CASES3D[geo]
    V = 24
    ang = angles_with_ties(V)
    vol, sino = shapes3d(opts, 2, V)
    x = np.random.default_rng(100 + OPS.index(op)).random(vol if op in VOLUME_IN else sino).astype(np.float32)
    options = torch.tensor(opts, dtype=torch.float32, device=cuda)
    angles = torch.from_numpy(ang).to(cuda)
    print(f"[gpu] _C.{NAMES3D[op]} {geo}", flush=True)
    got = getattr(_C(), NAMES3D[op])(torch.from_numpy(x).to(cuda).unsqueeze(1), options, angles)
    torch.cuda.synchronize()
    want = oracle.op3d(op, x.astype(np.float64), ang, opts)
    assert got.shape[1] == 1
    assert rel_l2(got.cpu().numpy()[:, 0], want) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES2D))
@pytest.mark.parametrize("name", sorted(NAMES2D.values()) + sorted(ALIASES2D))
def test_reference_named_2d_functions_vs_oracle(cuda, geo, name):
    op = ALIASES2D.get(name) or {v: k for k, v in NAMES2D.items()}[name]
    o2 = CASES2D[geo]
    W, H, Nu = int(o2[0]), int(o2[1]), int(o2[2])
    V = 40
    ang = angles_with_ties(V)
    rng = np.random.default_rng(200 + OPS.index(op))
    x = rng.random((2, H, W) if op in VOLUME_IN else (2, V, Nu)).astype(np.float32)
    options = torch.tensor(o2, dtype=torch.float32, device=cuda)
    print(f"[gpu] _C.{name} {geo}", flush=True)
    got = getattr(_C(), name)(torch.from_numpy(x).to(cuda).unsqueeze(1), options, torch.from_numpy(ang).to(cuda))
    torch.cuda.synchronize()
    want = oracle.op2d(op, x.astype(np.float64), ang, o2)
    assert rel_l2(got.cpu().numpy()[:, 0], want) <= 1e-5


def test_reference_named_functions_reuse_plans(cuda):
    C = _C()
    C.plan_cache_clear()
    opts = CASES3D["flat"]
    ang = angles_with_ties(16)
    options = torch.tensor(opts, dtype=torch.float32, device=cuda)
    angles = torch.from_numpy(ang).to(cuda)
    vol, _ = shapes3d(opts, 1, 16)
    x = torch.rand((1, 1) + vol[1:], device=cuda, generator=torch.Generator(device=cuda).manual_seed(5))
    first = C.forward_projection3d(x, options, angles)
    assert C.plan_cache_size() == 1
    for _ in range(3):  # same tensors: identity hit
        again = C.forward_projection3d(x, options, angles)
        assert torch.equal(again, first)
    # equal geometry in new tensors: content hit, still one plan
    again = C.forward_projection3d(x, options.clone(), angles.clone())
    assert torch.equal(again, first)
    assert C.plan_cache_size() == 1
    # an in-place change of the angles is seen (version counter), and makes a second plan
    angles.add_(0.25)
    moved = C.forward_projection3d(x, options, angles)
    assert C.plan_cache_size() == 2
    want = oracle.op3d("A", x.cpu().numpy()[:, 0].astype(np.float64), (ang + np.float32(0.25)).astype(np.float32), opts)
    assert rel_l2(moved.cpu().numpy()[:, 0], want) <= 1e-5


def test_reference_named_functions_run_on_current_stream(cuda):
    """With no stream argument the work goes to torch's current stream: results produced
    on a side stream are complete when that stream is synchronised, and a graph captured
    on a side stream (torch.cuda.graph) replays the same result."""
    C = _C()
    opts = CASES3D["arc"]
    ang = angles_with_ties(20)
    options = torch.tensor(opts, dtype=torch.float32, device=cuda)
    angles = torch.from_numpy(ang).to(cuda)
    vol, _ = shapes3d(opts, 1, 20)
    x = torch.rand((1, 1) + vol[1:], device=cuda, generator=torch.Generator(device=cuda).manual_seed(6))
    eager = C.forward_projection3d(x, options, angles)
    torch.cuda.synchronize()
    side = torch.cuda.Stream(cuda)
    side.wait_stream(torch.cuda.current_stream(cuda))
    with torch.cuda.stream(side):
        on_side = C.forward_projection3d(x, options, angles)
        on_side = on_side * 1.0
    side.synchronize()
    assert torch.equal(on_side, eager)
    # the plan exists (identity hit, no host copy), so the call can be captured
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        captured = C.forward_projection3d(x, options, angles)
    for rep in range(3):
        captured.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(captured, eager), f"replay {rep}"


# ------------------------------------------------------------ one-shot C-ABI via ctypes

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


# include/xrecon_amd.h: name and whether the first pointer argument is the volume
ONESHOT3D = {"A": ("xr_forward3d", True), "AT": ("xr_forward3d_t", True), "B": ("xr_backward3d", True),
             "BW": ("xr_weighted_backward3d", True), "BT": ("xr_backward3d_t", False),
             "BWT": ("xr_weighted_backward3d_t", False)}


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.xr_last_error.restype = ctypes.c_char_p
    return lib


@pytest.mark.parametrize("geo", sorted(CASES3D))
@pytest.mark.parametrize("op", OPS)
def test_oneshot_c_abi_3d_vs_oracle(cuda, lib, geo, op):
    o = CASES3D[geo]
    V, nb = 18, 2
    ang = angles_with_ties(V)
    vol, sino = shapes3d(o, nb, V)
    rng = np.random.default_rng(300 + OPS.index(op))
    x = rng.random(vol if op in VOLUME_IN else sino).astype(np.float32)
    dx = torch.from_numpy(x).to(cuda)
    out = torch.full(sino if op in VOLUME_IN else vol, float("nan"), device=cuda)
    v3 = _Vol3(int(o[0]), int(o[1]), int(o[2]), o[5], o[6], o[7], o[12], o[13], o[14])
    g3 = _Geo3(int(o[17]), int(o[3]), int(o[4]), o[8], o[9], o[15], o[16], o[10], o[11])
    name, image_first = ONESHOT3D[op]
    if op in VOLUME_IN:
        image, proj = dx, out
    else:
        image, proj = out, dx
    first, second = (image, proj) if image_first else (proj, image)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    angles_host = np.ascontiguousarray(ang)
    print(f"[gpu] {name} {geo}", flush=True)
    rc = getattr(lib, name)(ctypes.c_void_p(first.data_ptr()), ctypes.c_void_p(second.data_ptr()),
                            angles_host.ctypes.data_as(ctypes.c_void_p), ctypes.byref(v3), ctypes.byref(g3),
                            nb, V, ctypes.c_void_p(stream))
    assert rc == 0, lib.xr_last_error()
    torch.cuda.synchronize()
    want = oracle.op3d(op, x.astype(np.float64), ang, o)
    assert rel_l2(out.cpu().numpy(), want) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES2D))
@pytest.mark.parametrize("op", OPS)
def test_oneshot_c_abi_2d_vs_oracle(cuda, lib, geo, op):
    o2 = CASES2D[geo]
    W, H, Nu = int(o2[0]), int(o2[1]), int(o2[2])
    V, nb = 30, 2
    ang = angles_with_ties(V)
    rng = np.random.default_rng(400 + OPS.index(op))
    x = rng.random((nb, H, W) if op in VOLUME_IN else (nb, V, Nu)).astype(np.float32)
    out = torch.full((nb, V, Nu) if op in VOLUME_IN else (nb, H, W), float("nan"), device=cuda)
    dx = torch.from_numpy(x).to(cuda)
    v2 = _Vol2(W, H, o2[3], o2[4], o2[8], o2[9])
    g2 = _Geo2(int(o2[11]), Nu, o2[5], o2[10], o2[6], o2[7])
    angles_host = np.ascontiguousarray(ang)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    print(f"[gpu] xr_op2d {op} {geo}", flush=True)
    rc = lib.xr_op2d(OPS.index(op), ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                     angles_host.ctypes.data_as(ctypes.c_void_p), ctypes.byref(v2), ctypes.byref(g2), nb, V,
                     ctypes.c_void_p(stream))
    assert rc == 0, lib.xr_last_error()
    torch.cuda.synchronize()
    want = oracle.op2d(op, x.astype(np.float64), ang, o2)
    assert rel_l2(out.cpu().numpy(), want) <= 1e-5
