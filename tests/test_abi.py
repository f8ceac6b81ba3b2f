"""CPU: the C-ABI library loads and exports every function that include/xrecon_amd.h
declares (no compute calls: there is no GPU here), and the error path works without a
device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xrecon_amd.h")
LIB = os.path.join(ROOT, "xrecon_amd", "_lib", "libxrecon_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(xr_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    return ctypes.CDLL(LIB)


def test_header_declares_the_operator_surface():
    names = declared_functions()
    for required in ("xr_plan3d_create", "xr_plan3d_execute", "xr_plan3d_destroy", "xr_plan3d_workspace_bytes",
                     "xr_forward3d", "xr_forward3d_t", "xr_backward3d", "xr_weighted_backward3d",
                     "xr_backward3d_t", "xr_weighted_backward3d_t", "xr_op2d", "xr_last_error"):
        assert required in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared in include/xrecon_amd.h but not exported: {missing}"


def test_version_and_error_reporting(lib):
    lib.xr_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.xr_version()
    lib.xr_last_error.restype = ctypes.c_char_p

    class Vol(ctypes.Structure):
        _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)] + \
                   [(n, ctypes.c_float) for n in ("dx", "dy", "dz", "sx", "sy", "sz")]

    class Geo(ctypes.Structure):
        _fields_ = [("type", ctypes.c_int), ("nu", ctypes.c_int), ("nv", ctypes.c_int)] + \
                   [(n, ctypes.c_float) for n in ("du", "dv", "su", "sv", "sod", "sdd")]

    plan = ctypes.c_void_p()
    ang = (ctypes.c_float * 2)(0.0, 1.0)
    # invalid geometry type is rejected on the host, before any device call
    rc = lib.xr_plan3d_create(ctypes.byref(Vol(4, 4, 4, 1, 1, 1, 0, 0, 0)),
                              ctypes.byref(Geo(7, 8, 4, 1, 1, 0, 0, 500, 1000)), ang, 2, 0, ctypes.byref(plan))
    assert rc == -1
    assert b"scan type" in lib.xr_last_error()
    # negative spacing is rejected
    rc = lib.xr_plan3d_create(ctypes.byref(Vol(4, 4, 4, -1, 1, 1, 0, 0, 0)),
                              ctypes.byref(Geo(1, 8, 4, 1, 1, 0, 0, 500, 1000)), ang, 2, 0, ctypes.byref(plan))
    assert rc == -1
    assert b"spacing" in lib.xr_last_error()
    # cone geometry without distances is rejected
    rc = lib.xr_plan3d_create(ctypes.byref(Vol(4, 4, 4, 1, 1, 1, 0, 0, 0)),
                              ctypes.byref(Geo(1, 8, 4, 1, 1, 0, 0, 0, 0)), ang, 2, 0, ctypes.byref(plan))
    assert rc == -1
    lib.xr_plan3d_workspace_bytes.restype = ctypes.c_size_t
    assert lib.xr_plan3d_workspace_bytes(None, 0, 1) == 0
