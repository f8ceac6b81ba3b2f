"""Writes tests/golden/nnz.json: the exact number of (voxel, view, detector-cell) pairs
with positive footprint overlap for the benchmark workloads, counted by the FP64 oracle
(oracle/dd_oracle.c, counting mode).  bench.py reads these counts to turn kernel times
into algorithmic FLOP rates (F = 4 * nnz per operator); tests/test_oracle.py checks the
counting mode against a brute-force count on small geometries.

    python tests/golden/make_nnz.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

WORKLOADS = {
    # 256x256x64 @ 1 mm, flat 512x64 @ 1 mm, SOD 500, SDD 1000, 720 views over 2 pi
    "cone_flat_256x256x64_512x64_720v": (
        [256, 256, 64, 512, 64, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 0, 0, 0, 0, 0, 1], 720),
}


def main():
    out = {}
    for name, (opts, V) in WORKLOADS.items():
        ang = (np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V)).astype(np.float32)
        t = time.time()
        n = oracle.count3d(ang, opts)
        out[name] = {"options": opts, "num_view": V, "nnz": n,
                     "voxel_views": int(opts[0] * opts[1] * opts[2] * V), "seconds": round(time.time() - t, 1)}
        print(name, out[name], flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "nnz.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
