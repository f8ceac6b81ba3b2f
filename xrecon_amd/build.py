"""In-tree build of the native parts of xrecon_amd.

* ``xrecon_amd/_lib/libxrecon_amd.so`` — HIP kernels, FP64 plan builder and the C-ABI
  (``include/xrecon_amd.h``), compiled by ``hipcc --offload-arch=gfx950``.
* ``xrecon_amd/_C*.so`` — the PyTorch extension (``csrc/torch_binding.cpp``) that
  validates tensors and calls the C-ABI; compiled against torch's headers and linked to
  the library above with an ``$ORIGIN`` rpath.

Run ``python -m xrecon_amd.build`` (or ``__graft_entry__.build()``).  Both targets are
rebuilt on every call and replaced atomically.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libxrecon_amd.so")
ARCH = os.environ.get("XRECON_AMD_ARCH", "gfx950")

HIP_SOURCES = ["dd_kernels.hip", "batch.hip", "noalloc.hip", "sysmat_gpu.hip", "plan.cpp", "capi.cpp", "cpu_exec.cpp", "sysmat.cpp"]
# Packed-FP32 instructions are disabled for all device code (they have produced silent
# wrong results on this hardware generation in some forms).  No implicit contraction of
# a * b + c: every fused multiply-add in the kernels is an explicit fmaf, so results do not
# depend on how a template instance or launch shape happened to be scheduled (the tile
# shapes and the batched path are bitwise equal to the plain one), and the host emulator
# evaluates the same expressions.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in res.stdout.splitlines() if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def ext_path() -> str:
    return os.path.join(HERE, "_C" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_library(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", *HIP_FLAGS, "-o", tmp] + [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run(cmd, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_extension(verbose: bool = False) -> str:
    import torch
    from torch.utils import cpp_extension as ce

    out = ext_path()
    tmp = out + ".tmp"
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    # torch's HIP headers (c10/hip: the current stream of a device) need the ROCm include
    # directory and the AMD platform macro; the HIP runtime itself comes through c10_hip.
    inc = [f"-I{p}" for p in ce.include_paths(device_type="cuda")] + [f"-I{sysconfig.get_paths()['include']}"]
    libs = [f"-L{p}" for p in ce.library_paths()]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-w", "-D__HIP_PLATFORM_AMD__=1",
           "-DTORCH_EXTENSION_NAME=_C", "-DTORCH_API_INCLUDE_EXTENSION_H", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           *inc, os.path.join(CSRC, "torch_binding.cpp"), "-o", tmp,
           f"-L{LIBDIR}", "-lxrecon_amd", *libs, "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python",
           "-Wl,-rpath,$ORIGIN/_lib"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_extension(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, "and", ext_path())
