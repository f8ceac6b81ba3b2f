"""ORACLE — test infrastructure, not product code.

Python access to the two CPU checkers built by ``oracle/Makefile``:

* ``op3d`` / ``count3d`` / ``view3d`` — the FP64 restatement of the distance-driven
  discretisation (``oracle/dd_oracle.c``), with every formula cited there against the
  reference's ``include/xrmath.h``, ``include/cpu3d_common.h`` and ``src/cpu/*.cpp``.
* ``ref_op3d`` / ``ref_op2d`` / ``ref_sm3d`` / ``ref_sm2d`` — the reference's own CPU
  kernels (``src/cpu/*.cpp``), compiled from their sources into
  ``oracle/_ref/libxrecon_ref.so`` and called through ``oracle/ref_shim.cpp``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` (its CPU-baseline leg and
its nnz count) import this package, and only to check or to time the CPU baseline.

Geometry is passed as the reference's 18-entry 3D options vector
(``xrecon/projector/projector3d.py:98-107``) or 12-entry 2D vector
(``xrecon/projector/projector.py:106-110``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE_SO = os.path.join(_HERE, "libdd_oracle.so")
_REF_SO = os.path.join(_HERE, "_ref", "libxrecon_ref.so")

OPS = {"A": 0, "AT": 1, "B": 2, "BW": 3, "BT": 4, "BWT": 5}
_VOLUME_IN = {"A", "BT", "BWT"}

_lib = None
_ref = None


def _load_oracle():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {_ORACLE_SO} (run `make -C oracle`)")
        lib = ctypes.CDLL(_ORACLE_SO)
        lib.oracle_op3d.restype = ctypes.c_int64
        lib.oracle_op3d.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_view3d.restype = ctypes.c_int
        lib.oracle_view3d.argtypes = [ctypes.c_float] + [ctypes.c_void_p] * 8
        lib.oracle_op2d.restype = ctypes.c_int
        lib.oracle_op2d.argtypes = lib.oracle_op3d.argtypes
        _lib = lib
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        if not os.path.exists(_REF_SO):
            raise RuntimeError(f"reference CPU library missing: {_REF_SO} (run `make -C oracle ref`)")
        lib = ctypes.CDLL(_REF_SO)
        for name in ("ref_op3d", "ref_op2d"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 7 + [ctypes.c_int, ctypes.c_int]
        for name in ("ref_sm3d", "ref_sm2d"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int64
            fn.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int]
        lib.ref_sm_fetch.restype = None
        lib.ref_sm_fetch.argtypes = [ctypes.c_void_p] * 3
        _ref = lib
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def split3d(options: Sequence[float]):
    """Options[18] -> (vn, vf, gn, gf) in the order ``dd_oracle.c`` / ``ref_shim.cpp`` read."""
    o = np.asarray(options, dtype=np.float32)
    vn = np.array([int(o[0]), int(o[1]), int(o[2])], dtype=np.int32)
    vf = np.array([o[5], o[6], o[7], o[12], o[13], o[14]], dtype=np.float32)
    gn = np.array([int(o[17]), int(o[3]), int(o[4])], dtype=np.int32)
    gf = np.array([o[8], o[9], o[15], o[16], o[10], o[11]], dtype=np.float32)
    return vn, vf, gn, gf


def split2d(options: Sequence[float]):
    """Options[12] = W, H, Nu, dx, dy, du, SOD, SDD, shx, shy, shu, scan_id."""
    o = np.asarray(options, dtype=np.float32)
    vn = np.array([int(o[0]), int(o[1])], dtype=np.int32)
    vf = np.array([o[3], o[4], o[8], o[9]], dtype=np.float32)
    gn = np.array([int(o[11]), int(o[2])], dtype=np.int32)
    gf = np.array([o[5], o[10], o[6], o[7]], dtype=np.float32)
    return vn, vf, gn, gf


def _shapes3d(options, nbatch, nview):
    o = [float(v) for v in options]
    W, H, D, Nu, Nv = (int(v) for v in o[:5])
    return (nbatch, D, H, W), (nbatch, nview, Nv, Nu)


def op3d(op: str, x: np.ndarray, angles: np.ndarray, options: Sequence[float], nthreads: int | None = None) -> np.ndarray:
    """FP64 oracle for one 3D operator.  ``x`` is [nb, D, H, W] for A/BT/BWT and
    [nb, V, Nv, Nu] otherwise; the result is float64 of the other shape."""
    lib = _load_oracle()
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    nview = angles.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float64)
    nbatch = x.shape[0]
    vol_shape, sino_shape = _shapes3d(options, nbatch, nview)
    in_shape, out_shape = (vol_shape, sino_shape) if op in _VOLUME_IN else (sino_shape, vol_shape)
    if tuple(x.shape) != in_shape:
        raise ValueError(f"oracle {op}: expected input {in_shape}, got {x.shape}")
    out = np.zeros(out_shape, dtype=np.float64)
    vn, vf, gn, gf = split3d(options)
    lib.oracle_op3d(OPS[op], _ptr(x), _ptr(out), _ptr(angles), _ptr(vn), _ptr(vf), _ptr(gn), _ptr(gf),
                    nbatch, nview, nthreads or min(16, os.cpu_count() or 1))
    return out


def op2d(op: str, x: np.ndarray, angles: np.ndarray, options: Sequence[float], nthreads: int | None = None) -> np.ndarray:
    """FP64 oracle for one 2D operator: image [nb, H, W] <-> sinogram [nb, V, Nu]."""
    lib = _load_oracle()
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    nview = angles.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float64)
    nbatch = x.shape[0]
    o = [float(v) for v in options]
    W, H, Nu = int(o[0]), int(o[1]), int(o[2])
    in_shape, out_shape = ((nbatch, H, W), (nbatch, nview, Nu))
    if op not in _VOLUME_IN:
        in_shape, out_shape = out_shape, in_shape
    if tuple(x.shape) != in_shape:
        raise ValueError(f"oracle 2D {op}: expected input {in_shape}, got {x.shape}")
    out = np.zeros(out_shape, dtype=np.float64)
    vn, vf, gn, gf = split2d(options)
    lib.oracle_op2d(OPS[op], _ptr(x), _ptr(out), _ptr(angles), _ptr(vn), _ptr(vf), _ptr(gn), _ptr(gf),
                    nbatch, nview, nthreads or min(16, os.cpu_count() or 1))
    return out


def count3d(angles: np.ndarray, options: Sequence[float], nthreads: int | None = None) -> int:
    """Exact number of (voxel, view, cell) pairs with positive footprint overlap."""
    lib = _load_oracle()
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    vn, vf, gn, gf = split3d(options)
    dummy = np.zeros(1, dtype=np.float64)
    return int(lib.oracle_op3d(6, _ptr(dummy), _ptr(dummy), _ptr(angles), _ptr(vn), _ptr(vf), _ptr(gn),
                               _ptr(gf), 1, angles.shape[0], nthreads or min(16, os.cpu_count() or 1)))


def view3d(angle: float, options: Sequence[float]):
    """FP64 per-view geometry: (xdrv, keep, (sx, sy), d[Nu+1], sigma[Nu])."""
    lib = _load_oracle()
    vn, vf, gn, gf = split3d(options)
    nu = int(gn[1])
    flags = np.zeros(2, dtype=np.int32)
    src = np.zeros(2, dtype=np.float64)
    d = np.zeros(nu + 1, dtype=np.float64)
    sig = np.zeros(nu, dtype=np.float64)
    lib.oracle_view3d(ctypes.c_float(angle), _ptr(vn), _ptr(vf), _ptr(gn), _ptr(gf), _ptr(flags), _ptr(src),
                      _ptr(d), _ptr(sig))
    return bool(flags[0]), bool(flags[1]), (src[0], src[1]), d, sig


def ref_op3d(op: str, x: np.ndarray, angles: np.ndarray, options: Sequence[float]) -> np.ndarray:
    """The reference's own CPU kernel for one 3D operator (float32 in, float32 out)."""
    lib = _load_ref()
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    nview = angles.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float32)
    nbatch = x.shape[0]
    vol_shape, sino_shape = _shapes3d(options, nbatch, nview)
    if op in _VOLUME_IN:
        image, proj = x, np.zeros(sino_shape, dtype=np.float32)
    else:
        image, proj = np.zeros(vol_shape, dtype=np.float32), x
    vn, vf, gn, gf = split3d(options)
    rc = lib.ref_op3d(OPS[op], _ptr(image), _ptr(proj), _ptr(angles), _ptr(vn), _ptr(vf), _ptr(gn), _ptr(gf),
                      nbatch, nview)
    if rc != 0:
        raise RuntimeError(f"ref_op3d({op}) failed: {rc}")
    return proj if op in _VOLUME_IN else image


def ref_op2d(op: str, x: np.ndarray, angles: np.ndarray, options: Sequence[float]) -> np.ndarray:
    """The reference's own 2D CPU kernel: image [nb, H, W] <-> sinogram [nb, V, Nu]."""
    lib = _load_ref()
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    nview = angles.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float32)
    nbatch = x.shape[0]
    o = [float(v) for v in options]
    W, H, Nu = int(o[0]), int(o[1]), int(o[2])
    if op in _VOLUME_IN:
        image, proj = x, np.zeros((nbatch, nview, Nu), dtype=np.float32)
    else:
        image, proj = np.zeros((nbatch, H, W), dtype=np.float32), x
    vn, vf, gn, gf = split2d(options)
    rc = lib.ref_op2d(OPS[op], _ptr(image), _ptr(proj), _ptr(angles), _ptr(vn), _ptr(vf), _ptr(gn), _ptr(gf),
                      nbatch, nview)
    if rc != 0:
        raise RuntimeError(f"ref_op2d({op}) failed: {rc}")
    return proj if op in _VOLUME_IN else image


def _ref_sm(fn_name, angles, vn, vf, gn, gf):
    lib = _load_ref()
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    n = getattr(lib, fn_name)(_ptr(angles), _ptr(vn), _ptr(vf), _ptr(gn), _ptr(gf), angles.shape[0])
    rows = np.zeros(n, dtype=np.int64)
    cols = np.zeros(n, dtype=np.int64)
    vals = np.zeros(n, dtype=np.float32)
    lib.ref_sm_fetch(_ptr(rows), _ptr(cols), _ptr(vals))
    return rows, cols, vals


def ref_sm3d(angles: np.ndarray, options: Sequence[float]):
    """Reference system-matrix triplets (rows, cols, vals), before coalescing."""
    return _ref_sm("ref_sm3d", angles, *split3d(options))


def ref_sm2d(angles: np.ndarray, options: Sequence[float]):
    return _ref_sm("ref_sm2d", angles, *split2d(options))
