"""Writes tests/golden/ref_cpu_fixtures.npz: input/output vectors of the reference's own
CPU kernels (src/cpu/*.cpp, compiled from the reference sources into
oracle/_ref/libxrecon_ref.so by oracle/Makefile) on small seeded cases.

Each case stores the options vector, the float32 angles, the float32 input and the
reference's float32 output, for every operator of every 3D and 2D geometry, plus the
system-matrix triplets of two tiny geometries.  The CPU tests check the FP64 oracle
against these vectors, so the oracle stays pinned to the reference where the reference
cannot run (the GPU box, CI without the reference tree).

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from helpers import opts2d, opts3d  # noqa: E402

OPS = ("A", "AT", "B", "BW", "BT", "BWT")
VOL_IN = ("A", "BT", "BWT")

CASES3D = {
    "flat": opts3d(16, 14, 6, 24, 6, 1.0, 1.1, 0.9, 1.0, 1.0, 500.0, 1000.0, 1, 0.3, -0.2, 0.1, 0.25, 0.1),
    "arc": opts3d(16, 14, 6, 24, 6, 1.0, 1.1, 0.9, 2.0 / 1000.0, 1.0, 500.0, 1000.0, 2, 0.3, -0.2, 0.1, 0.5 / 1000.0, 0.1),
    "par": opts3d(16, 14, 6, 24, 6, 1.0, 1.2, 1.0, 0.9, 1.1, 0.0, 0.0, 0, 0.3, -0.2, 0.1, 0.2, 0.1),
}
CASES2D = {
    "flat": opts2d(20, 18, 30, 1.0, 1.1, 0.5, 500.0, 1000.0, 1, 0.3, -0.2, 0.25),
    "arc": opts2d(20, 18, 30, 1.0, 1.1, 2 * np.arctan(1.0 / 2000.0), 500.0, 1000.0, 2, 0.3, -0.2, 0.0003),
    "par": opts2d(20, 18, 30, 1.0, 1.1, 0.9, 0.0, 0.0, 0, 0.3, -0.2, 0.25),
}


def angle_sets():
    V = 16
    full = (np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V)).astype(np.float32)  # includes 45-degree ties
    small = np.linspace(-0.6, 0.6, 9).astype(np.float32)  # |beta| small: FP32 angle sums stay accurate
    return {"full": full, "small": small}


def main():
    if not oracle.ref_available():
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    rng = np.random.default_rng(2024)
    out = {}
    for dim, cases in (("3d", CASES3D), ("2d", CASES2D)):
        for geo, o in cases.items():
            for aname, ang in angle_sets().items():
                W, H = int(o[0]), int(o[1])
                if dim == "3d":
                    D, Nu, Nv = int(o[2]), int(o[3]), int(o[4])
                    vol, sino = (1, D, H, W), (1, len(ang), Nv, Nu)
                else:
                    Nu = int(o[2])
                    vol, sino = (1, H, W), (1, len(ang), Nu)
                xv = rng.random(vol).astype(np.float32)
                xs = rng.random(sino).astype(np.float32)
                for op in OPS:
                    x = xv if op in VOL_IN else xs
                    y = (oracle.ref_op3d if dim == "3d" else oracle.ref_op2d)(op, x, ang, o)
                    key = f"{dim}_{geo}_{aname}_{op}"
                    out[key + "_options"] = np.asarray(o, dtype=np.float64)
                    out[key + "_angles"] = ang
                    out[key + "_in"] = x
                    out[key + "_out"] = y
    # system-matrix triplets (reference system_matrix3d.cpp / system_matrix2d.cpp)
    sm3 = opts3d(6, 5, 3, 8, 3, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 1, 0.2, 0.1, 0.0, 0.3, 0.1)
    ang3 = (np.arange(6, dtype=np.float32) * np.float32(2 * np.pi / 6)).astype(np.float32)
    r, c, v = oracle.ref_sm3d(ang3, sm3)
    out.update({"sm3d_options": np.asarray(sm3), "sm3d_angles": ang3, "sm3d_rows": r, "sm3d_cols": c, "sm3d_vals": v})
    sm2 = opts2d(8, 7, 10, 1.0, 1.0, 0.5, 500.0, 1000.0, 1, 0.2, 0.1, 0.3)
    ang2 = (np.arange(8, dtype=np.float32) * np.float32(2 * np.pi / 8)).astype(np.float32)
    r, c, v = oracle.ref_sm2d(ang2, sm2)
    out.update({"sm2d_options": np.asarray(sm2), "sm2d_angles": ang2, "sm2d_rows": r, "sm2d_cols": c, "sm2d_vals": v})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_cpu_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays", os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
