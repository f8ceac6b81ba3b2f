"""GPU: the allocation-free entry point xr_plan3d_execute_noalloc (workspace from the
library's static device scratch) gives bitwise the same results as xr_plan3d_execute with
a caller-provided workspace, through the C-ABI, on two calls in a row whose second input is
smaller than the first, for every operator; and it rejects an operator whose workspace
exceeds the scratch before launching anything."""
import ctypes
import os

import numpy as np
import pytest
import torch

from helpers import CASES3D, OPS, VOLUME_IN, angles_with_ties, opts3d, shapes3d

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "xrecon_amd", "_lib", "libxrecon_amd.so")
XR_EWORKSPACE = -4
SCRATCH_BYTES = 256 << 20


class _Vol3(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("dx", "dy", "dz", "shift_x", "shift_y", "shift_z")]


class _Geo3(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("nu", ctypes.c_int), ("nv", ctypes.c_int)] + \
               [(n, ctypes.c_float) for n in ("du", "dv", "shift_u", "shift_v", "sod", "sdd")]


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.xr_last_error.restype = ctypes.c_char_p
    lib.xr_plan3d_workspace_bytes.restype = ctypes.c_size_t
    lib.xr_plan3d_workspace_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.xr_plan3d_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p]
    lib.xr_plan3d_execute.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.xr_plan3d_execute_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int, ctypes.c_void_p]
    lib.xr_plan3d_destroy.argtypes = [ctypes.c_void_p]
    lib.xr_plan3d_destroy.restype = None
    return lib


def _create(lib, o, ang):
    v3 = _Vol3(int(o[0]), int(o[1]), int(o[2]), o[5], o[6], o[7], o[12], o[13], o[14])
    g3 = _Geo3(int(o[17]), int(o[3]), int(o[4]), o[8], o[9], o[15], o[16], o[10], o[11])
    ang = np.ascontiguousarray(ang, dtype=np.float32)
    plan = ctypes.c_void_p()
    rc = lib.xr_plan3d_create(ctypes.byref(v3), ctypes.byref(g3), ang.ctypes.data, ang.shape[0], 0,
                              ctypes.byref(plan))
    assert rc == 0, lib.xr_last_error()
    return plan


# first call: the larger case; second call: a smaller volume, detector, view set and batch
CALLS = {
    "flat": [(CASES3D["flat"], 40, 3),
             (opts3d(12, 10, 4, 20, 5, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 1, 0.1, 0.0, 0.0, 0.2, 0.0), 9, 1)],
    "arc": [(CASES3D["arc"], 36, 2),
            (opts3d(10, 14, 3, 16, 4, 1.0, 1.0, 1.0, 1.0 / 1000.0, 1.0, 500.0, 1000.0, 2), 7, 1)],
}


@pytest.mark.parametrize("geo", sorted(CALLS))
@pytest.mark.parametrize("op", OPS)
def test_noalloc_matches_execute_on_two_calls(cuda, lib, geo, op):
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    k = OPS.index(op)
    for call, (o, V, nb) in enumerate(CALLS[geo]):
        ang = angles_with_ties(V)
        plan = _create(lib, o, ang)
        try:
            vol, sino = shapes3d(o, nb, V)
            gen = torch.Generator(device=cuda).manual_seed(700 + 10 * k + call)
            x = torch.rand(vol if op in VOLUME_IN else sino, device=cuda, generator=gen)
            out_shape = sino if op in VOLUME_IN else vol
            ws_bytes = lib.xr_plan3d_workspace_bytes(plan, k, nb)
            assert ws_bytes <= SCRATCH_BYTES
            ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=cuda)
            want = torch.full(out_shape, float("nan"), device=cuda)
            got = torch.full(out_shape, float("nan"), device=cuda)
            print(f"[gpu] noalloc {geo} {op} call {call}: in {tuple(x.shape)}", flush=True)
            rc = lib.xr_plan3d_execute(plan, k, x.data_ptr(), want.data_ptr(), nb, ws.data_ptr(), ws_bytes, stream)
            assert rc == 0, lib.xr_last_error()
            rc = lib.xr_plan3d_execute_noalloc(plan, k, x.data_ptr(), got.data_ptr(), nb, stream)
            assert rc == 0, lib.xr_last_error()
            torch.cuda.synchronize()
            assert torch.isfinite(want).all()
            assert torch.equal(got, want), f"call {call}"
        finally:
            torch.cuda.synchronize()
            lib.xr_plan3d_destroy(plan)


def test_noalloc_rejects_workspace_beyond_scratch(cuda, lib):
    """A 512^2 x 128 volume: its z-prefix tables (both driving axes) need about 541 MB,
    more than the 256 MiB scratch; the call fails on the host and launches nothing."""
    o = opts3d(512, 512, 128, 64, 8, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 1)
    ang = np.array([0.0, 0.8, 1.6, 2.4], dtype=np.float32)
    plan = _create(lib, o, ang)
    try:
        assert lib.xr_plan3d_workspace_bytes(plan, 0, 1) > SCRATCH_BYTES
        out = torch.zeros((1, 4, 8, 64), device=cuda)
        vol = torch.zeros((1, 128, 512, 512), device=cuda)
        stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
        rc = lib.xr_plan3d_execute_noalloc(plan, 0, vol.data_ptr(), out.data_ptr(), 1, stream)
        assert rc == XR_EWORKSPACE
        assert b"static scratch" in lib.xr_last_error()
        torch.cuda.synchronize()
        assert torch.count_nonzero(out).item() == 0
    finally:
        lib.xr_plan3d_destroy(plan)
