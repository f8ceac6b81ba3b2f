"""Writes tests/golden/nnz.json: the exact number of (voxel, view, detector-cell) pairs
with positive footprint overlap for the benchmark workloads, counted by the FP64 oracle
(oracle/dd_oracle.c, counting mode).  bench.py reads these counts to turn kernel times
into algorithmic FLOP rates (F = 4 * nnz per operator); tests/test_oracle.py checks the
counting mode against a brute-force count on small geometries.

    python tests/golden/make_nnz.py [--all]   (--all recounts entries already present)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

# name: (options, views over 2 pi, views counted).  When fewer views are counted than the
# workload has, they are evenly spaced (every V/n-th view of the uniform grid) and the
# count is scaled by V/n: an estimate, recorded as such ("sampled_views").
WORKLOADS = {
    # 256x256x64 @ 1 mm, flat 512x64 @ 1 mm, SOD 500, SDD 1000, 720 views over 2 pi
    "cone_flat_256x256x64_512x64_720v": (
        [256, 256, 64, 512, 64, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 0, 0, 0, 0, 0, 1], 720, 720),
    # 512x512x256 @ 0.7 mm, arc 736x128 (1.2 mm / 1.0 mm), SOD 595, SDD 1085.6, 1152 views
    "cone_arc_512x512x256_736x128_1152v": (
        [512, 512, 256, 736, 128, 0.7, 0.7, 0.7, float(np.float32(1.2 / 1085.6)), 1.0, 595.0, 1085.6,
         0, 0, 0, 0, 0, 2], 1152, 36),
}


def main():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nnz.json")
    out = {}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    redo = "--all" in sys.argv
    for name, (opts, V, n_count) in WORKLOADS.items():
        if name in out and not redo:
            continue
        ang = (np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V)).astype(np.float32)
        sel = ang[:: V // n_count][:n_count]
        t = time.time()
        n = oracle.count3d(np.ascontiguousarray(sel), opts)
        ent = {"options": opts, "num_view": V, "nnz": int(round(n * V / n_count)),
               "voxel_views": int(opts[0] * opts[1] * opts[2] * V), "seconds": round(time.time() - t, 1)}
        if n_count < V:
            ent["sampled_views"] = n_count
            ent["nnz_sampled"] = n
        out[name] = ent
        print(name, ent, flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
