"""Shared test helpers: geometry cases, error norms, and the CPU emulator of the kernel
bodies (tests/native/libxr_emu.so)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_SO = os.path.join(HERE, "native", "libxr_emu.so")

OPS = ("A", "AT", "B", "BW", "BT", "BWT")
VOLUME_IN = ("A", "BT", "BWT")
OP_ID = {"A": 0, "AT": 1, "B": 2, "BW": 3, "BT": 4, "BWT": 5}


def opts3d(W, H, D, Nu, Nv, dx, dy, dz, du, dv, SOD, SDD, scan, shx=0.0, shy=0.0, shz=0.0, shu=0.0, shv=0.0):
    """The reference's 18-entry 3D options vector (projector3d.py:98-107)."""
    return [W, H, D, Nu, Nv, dx, dy, dz, du, dv, SOD, SDD, shx, shy, shz, shu, shv, scan]


def opts2d(W, H, Nu, dx, dy, du, SOD, SDD, scan, shx=0.0, shy=0.0, shu=0.0):
    """The reference's 12-entry 2D options vector (projector.py:106-110)."""
    return [W, H, Nu, dx, dy, du, SOD, SDD, shx, shy, shu, scan]


# Small but non-trivial geometries: non-square pixels, shifted volume and detector,
# enough rows and columns to cross several tile boundaries.
CASES3D = {
    "flat": opts3d(32, 28, 10, 48, 12, 1.0, 1.1, 0.9, 1.0, 1.0, 500.0, 1000.0, 1, 0.3, -0.2, 0.1, 0.25, 0.1),
    "arc": opts3d(32, 28, 10, 48, 12, 1.0, 1.1, 0.9, 1.0 / 1000.0, 1.0, 500.0, 1000.0, 2, 0.3, -0.2, 0.1,
                  0.25 / 1000.0, 0.1),
    "par": opts3d(24, 20, 6, 40, 8, 1.0, 1.2, 1.0, 0.9, 1.1, 0.0, 0.0, 0, 0.3, -0.2, 0.1, 0.2, 0.1),
}


# 2D cases: odd sizes, non-square pixels, shifted image and detector.
CASES2D = {
    "flat": opts2d(33, 17, 45, 1.0, 1.1, 0.5, 500.0, 1000.0, 1, 0.3, -0.2, 0.25),
    "arc": opts2d(33, 17, 45, 1.0, 1.1, 2 * np.arctan(1.0 / 2000.0), 500.0, 1000.0, 2, 0.3, -0.2, 0.0003),
    "par": opts2d(33, 17, 45, 1.0, 1.1, 0.9, 0.0, 0.0, 0, 0.3, -0.2, 0.25),
}


def angles_with_ties(V: int, offset: float = 0.01) -> np.ndarray:
    """V float32 angles over 2 pi with the 45-degree tie views inserted exactly as the
    float32 grid k*2pi/8 produces them."""
    a = (np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V) + np.float32(offset)).astype(np.float32)
    for q, k in enumerate((1, 3, 5, 7)):
        a[(q * V) // 4 + 1] = np.float32(k * 2 * np.pi / 8)
    return a


def uniform_angles(V: int) -> np.ndarray:
    return (np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V)).astype(np.float32)


def shapes3d(opts, nb, V):
    W, H, D, Nu, Nv = (int(v) for v in opts[:5])
    return (nb, D, H, W), (nb, V, Nv, Nu)


def correlate_fp64(sino, filt):
    """out[..., u] = sum_s sino[..., s] * filt[s - u + Nu - 1] with zero padding, in FP64:
    the reference's conv over the detector axis with padding Nu - 1 (projector3d.py:190,
    projector.py:201), written directly."""
    sino = np.asarray(sino, dtype=np.float64)
    filt = np.asarray(filt, dtype=np.float64).ravel()
    nu = sino.shape[-1]
    out = np.zeros(sino.shape, dtype=np.float64)
    for u in range(nu):
        out[..., u] = sino @ filt[nu - 1 - u: 2 * nu - 1 - u]
    return out


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


_emu = None


def emu_available() -> bool:
    return os.path.exists(EMU_SO)


def _load_emu():
    global _emu
    if _emu is None:
        lib = ctypes.CDLL(EMU_SO)
        lib.emu_op3d.restype = ctypes.c_int
        lib.emu_op3d.argtypes = ([ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_void_p,
                                                                            ctypes.c_void_p] + [ctypes.c_int] * 4)
        lib.emu_error.restype = ctypes.c_char_p
        _emu = lib
    return _emu


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def emu_op3d(op, x, angles, opts, R=0, KC=0, is2d=False, nthreads=None, cell_variant=0, voxel_variant=0):
    """The product's kernel bodies run on the CPU (float32 in/out).  cell_variant: 0 = each
    thread reading the table (the direct launch), 1 = general path with 32-bit offsets,
    2 = general path with 64-bit offsets, 3 = the LDS-staged workgroups (tiles of 4 rows
    and more).  voxel_variant: 0 = the workgroup path with the
    per-view block staged (as launched), 1 = every lookup from the table."""
    lib = _load_emu()
    lib.emu_set_cell_variant(int(cell_variant))
    lib.emu_set_voxel_variant(int(voxel_variant))
    vn, vf, gn, gf = oracle.split3d(opts)
    if is2d:
        gn = gn.copy()
        gn[0] |= 0x100
    x = np.ascontiguousarray(x, dtype=np.float32)
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    nb = x.shape[0]
    vol, sino = shapes3d(opts, nb, angles.shape[0])
    out = np.zeros(sino if op in VOLUME_IN else vol, dtype=np.float32)
    rc = lib.emu_op3d(OP_ID[op], _p(vn), _p(vf), _p(gn), _p(gf), _p(angles), angles.shape[0], _p(x), _p(out), nb,
                      R, KC, nthreads or min(8, os.cpu_count() or 1))
    if rc != 0:
        raise RuntimeError(lib.emu_error().decode())
    return out


def emu_cell_lines(reset=True):
    """The staged cell gather's counts since the last reset (emulator variant 3):
    (thread, line) pairs served from an LDS block, lookups outside their block (must be
    0), workgroups staged, workgroups direct."""
    lib = _load_emu()
    counts = np.zeros(4, dtype=np.int64)
    lib.emu_cell_lines(counts.ctypes.data_as(ctypes.c_void_p), int(reset))
    return counts


def emu_cell_box_stats(angles, opts, R2, R3):
    """Block statistics of the staged cell gather over a plan (tiles of R2 rows for the
    two-pixel class, R3 for the three-pixel one): [blocks, blocks over a slot, (workgroup,
    row tile) pairs staged, pairs direct, entries staged, largest block, blocks over 1024
    entries]."""
    lib = _load_emu()
    vn, vf, gn, gf = oracle.split3d(opts)
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    st = np.zeros(7, dtype=np.int64)
    rc = lib.emu_cell_box_stats(_p(vn), _p(vf), _p(gn), _p(gf), _p(angles), angles.shape[0], int(R2), int(R3),
                                st.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError(lib.emu_error().decode())
    return st


def emu_fdk3d(y, angles, opts, filt, weight, is2d=False, KC=0, nthreads=None):
    """The fused FDK pass on the CPU: sinogram [nb, V, Nv, Nu] -> volume [nb, D, H, W]."""
    lib = _load_emu()
    vn, vf, gn, gf = oracle.split3d(opts)
    if is2d:
        gn = gn.copy()
        gn[0] |= 0x100
    y = np.ascontiguousarray(y, dtype=np.float32)
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    filt = np.ascontiguousarray(filt, dtype=np.float32).ravel()
    weight = np.ascontiguousarray(weight, dtype=np.float32).ravel()
    nb = y.shape[0]
    vol, _ = shapes3d(opts, nb, angles.shape[0])
    out = np.zeros(vol, dtype=np.float32)
    rc = lib.emu_fdk3d(_p(vn), _p(vf), _p(gn), _p(gf), _p(angles), angles.shape[0], _p(y), _p(filt), _p(weight),
                       _p(out), nb, KC, nthreads or min(8, os.cpu_count() or 1))
    if rc != 0:
        raise RuntimeError(lib.emu_error().decode())
    return out


def emu_view_spans(angles, opts, groups=False):
    """Span class per view (2 / 3: fixed-span cell path, 0: general loop); with
    groups=True also the class per (view, 64-column group), which the kernels use."""
    lib = _load_emu()
    vn, vf, gn, gf = oracle.split3d(opts)
    angles = np.ascontiguousarray(angles, dtype=np.float32)
    spans = np.zeros(angles.shape[0], dtype=np.int32)
    ncg = -(-int(opts[3]) // 64)
    gsp = np.zeros((angles.shape[0], ncg), dtype=np.int32)
    rc = lib.emu_view_spans(_p(vn), _p(vf), _p(gn), _p(gf), _p(angles), angles.shape[0], _p(spans), _p(gsp))
    if rc != 0:
        raise RuntimeError(lib.emu_error().decode())
    return (spans, gsp) if groups else spans


def opts2d_as_3d(o2):
    """2D options -> the one-slice 3D options the kernels use (with the 2D flag set by
    emu_op3d(is2d=True)): W, H, 1, Nu, 1, dx, dy, 1, du, 1, SOD, SDD, shx, shy, 0, shu, 0, type."""
    W, H, Nu, dx, dy, du, SOD, SDD, shx, shy, shu, scan = o2
    return [W, H, 1, Nu, 1, dx, dy, 1.0, du, 1.0, SOD, SDD, shx, shy, 0.0, shu, 0.0, scan]
