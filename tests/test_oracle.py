"""CPU: pins the FP64 oracle.

* against golden vectors of the reference's own CPU kernels (tests/golden/
  ref_cpu_fixtures.npz, written by tests/golden/make_golden.py from oracle/_ref), and
  live against oracle/_ref where it is built;
* against what the mathematics fixes, chosen so that a plausible mistake fails one of
  them: exact adjointness (transposition/indexing), the chord length of a uniform slab
  for central rays (the Delta/(W H cos) normalisation and the cosine), the 90-degree
  rotation symmetry (x-driven vs y-driven branches and the flags), brute-force overlap
  counts (the nnz used for FLOP accounting), and the reference's system matrix
  multiplied by a volume (the detector-driven traversal vs the oracle's voxel-driven
  enumeration);
* and a planted mistake: a perturbed oracle geometry must fail the golden comparison.
"""
import os

import numpy as np
import pytest

import oracle
from helpers import opts3d, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "ref_cpu_fixtures.npz")

OPS = ("A", "AT", "B", "BW", "BT", "BWT")


def _fixtures():
    return np.load(FIX)


def _tol(geo, aset, op):
    # The reference evaluates arc column boundaries as tanf(beta + gamma) in FP32, whose
    # error grows with |beta| (measured: ~1e-6 at beta ~ 0, ~1e-4 at beta ~ 2 pi).
    if geo == "arc" and aset == "full":
        return 3e-4
    if geo == "arc":
        return 5e-5
    return 3e-6


@pytest.mark.parametrize("dim", ["3d", "2d"])
@pytest.mark.parametrize("geo", ["flat", "arc", "par"])
@pytest.mark.parametrize("aset", ["full", "small"])
def test_oracle_matches_reference_golden(dim, geo, aset):
    f = _fixtures()
    for op in OPS:
        key = f"{dim}_{geo}_{aset}_{op}"
        opts, ang, x, want = (f[key + s] for s in ("_options", "_angles", "_in", "_out"))
        run = oracle.op3d if dim == "3d" else oracle.op2d
        got = run(op, x.astype(np.float64), ang, list(opts))
        err = rel_l2(got, want)
        assert err <= _tol(geo, aset, op), f"{key}: relL2 {err:.3e}"
        if geo == "arc" and op in ("A", "BT", "BWT"):
            # at beta = 0 the reference's FP32 angle sums are exact to a few ulp
            v0 = int(np.argmin(np.abs(ang)))
            assert abs(float(ang[v0])) < 1e-6
            assert rel_l2(got[:, v0], want[:, v0]) <= 3e-6, f"{key} at beta=0"


def test_planted_mistake_is_caught():
    """A 0.1% error in the source distance (a plausible units/scale slip) must fail."""
    f = _fixtures()
    key = "3d_flat_full_A"
    opts = list(f[key + "_options"])
    opts[10] *= 1.001
    got = oracle.op3d("A", f[key + "_in"].astype(np.float64), f[key + "_angles"], opts)
    assert rel_l2(got, f[key + "_out"]) > 1e-4


@pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU library not built")
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_matches_reference_live(seed):
    rng = np.random.default_rng(seed)
    W, H, D = (int(v) for v in rng.integers(8, 20, size=3))
    Nu, Nv = int(rng.integers(10, 30)), int(rng.integers(3, 9))
    o = opts3d(W, H, D, Nu, Nv, *rng.uniform(0.7, 1.3, size=3), float(rng.uniform(0.8, 1.6)),
               float(rng.uniform(0.8, 1.5)), 400.0, 900.0, 1, *rng.uniform(-0.5, 0.5, size=3),
               float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)))
    ang = np.sort(rng.uniform(0, 2 * np.pi, size=12)).astype(np.float32)
    x = rng.random((1, D, H, W)).astype(np.float32)
    for op in ("A", "BT"):
        assert rel_l2(oracle.ref_op3d(op, x, ang, o), oracle.op3d(op, x.astype(np.float64), ang, o)) < 3e-6


@pytest.mark.parametrize("dim", ["3d", "2d"])
@pytest.mark.parametrize("geo", ["flat", "arc", "par"])
def test_oracle_adjoint_pairs(dim, geo):
    f = _fixtures()
    rng = np.random.default_rng(5)
    run = oracle.op3d if dim == "3d" else oracle.op2d
    for fwd, adj in (("A", "AT"), ("BT", "B"), ("BWT", "BW")):
        key = f"{dim}_{geo}_full_{fwd}"
        opts, ang, x = list(f[key + "_options"]), f[key + "_angles"], f[key + "_in"].astype(np.float64)
        y = rng.random(f[key + "_out"].shape)
        lhs = np.sum(run(fwd, x, ang, opts) * y)
        rhs = np.sum(x * run(adj, y, ang, opts))
        assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def test_uniform_slab_chord_length():
    """A of a uniform volume along a central ray = slab thickness / cos(theta), with the
    cosine computed independently from the source and cell-centre positions."""
    W = H = 40
    D, Nu, Nv, sod, sdd = 40, 8, 6, 500.0, 1000.0
    o = opts3d(W, H, D, Nu, Nv, 1.0, 1.0, 1.0, 1.0, 1.0, sod, sdd, 1)
    ang = np.array([0.0], np.float32)
    y = oracle.op3d("A", np.ones((1, D, H, W)), ang, o)[0, 0]
    mag = sod / sdd
    for r in range(Nv):
        for c in range(Nu):
            u = (c + 0.5 - Nu / 2) * 1.0 * mag  # virtual (iso-plane) cell centre
            v = (r + 0.5 - Nv / 2) * 1.0 * mag
            cos_t = sod / np.sqrt(u * u + v * v + sod * sod)
            assert abs(y[r, c] - H * 1.0 / cos_t) <= 1e-9 * y[r, c]


def test_rotation_symmetry_crosses_driving_axis():
    o = opts3d(12, 12, 4, 20, 4, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 1)
    b = np.float32(0.3)  # x-driven; b - pi/2 is y-driven
    x = np.random.default_rng(0).random((1, 4, 12, 12))
    xr = np.rot90(x, k=1, axes=(2, 3)).copy()
    y_rot = oracle.op3d("A", xr, np.array([b], np.float32), o)
    y = oracle.op3d("A", x, np.array([np.float32(b - np.pi / 2)], np.float32), o)
    assert rel_l2(y_rot, y) < 1e-6


def test_count_matches_brute_force():
    """count3d (the nnz behind the FLOP accounting) against an explicit enumeration of
    every (voxel, view, cell) triple with positive overlap, from the oracle's own FP64
    per-view geometry."""
    o = opts3d(6, 5, 3, 9, 4, 1.0, 1.2, 0.8, 1.0, 1.0, 500.0, 1000.0, 1, 0.2, -0.1, 0.1, 0.3, 0.2)
    ang = np.array([0.1, 0.9, 2.2, 3.9, np.float32(np.pi / 4)], np.float32)
    n = 0
    W, H, D, Nu, Nv = 6, 5, 3, 9, 4
    dx, dy, dz, sod, sdd = 1.0, 1.2, 0.8, 500.0, 1000.0
    shx, shy, shz, shv = 0.2, -0.1, 0.1, 0.2
    for a in ang:
        xdrv, keep, (sx, sy), d, sig = oracle.view3d(float(a), o)
        for k in range(D):
            for j in range(H):
                for i in range(W):
                    px = (i - W / 2 + 0.5) * dx + shx
                    py = (H / 2 - j - 0.5) * dy + shy
                    pz = (k - D / 2 + 0.5) * dz + shz
                    if xdrv:
                        amap = lambda x_, y_: (sx * y_ - sy * x_) / (y_ - sy)  # noqa: E731
                        a0, a1 = amap(px - dx / 2, py), amap(px + dx / 2, py)
                        zeta = sy / (sy - py)
                    else:
                        amap = lambda x_, y_: (sy * x_ - sx * y_) / (x_ - sx)  # noqa: E731
                        a0, a1 = amap(px, py - dy / 2), amap(px, py + dy / 2)
                        zeta = sx / (sx - px)
                    z0, z1 = zeta * (pz - dz / 2), zeta * (pz + dz / 2)
                    for c in range(Nu):
                        if min(a1, d[c + 1]) - max(a0, d[c]) <= 0:
                            continue
                        for r in range(Nv):
                            v0 = ((r - Nv / 2) * 1.0 + shv) * sod / sdd
                            v1 = v0 + 1.0 * sod / sdd
                            if min(z1, sig[c] * v1) - max(z0, sig[c] * v0) > 0:
                                n += 1
    assert oracle.count3d(ang, o) == n


def test_system_matrix_times_volume_equals_projection():
    """The reference's own system matrix (its detector-driven traversal, emitted as
    triplets) applied to a volume equals the oracle's voxel-driven A, as
    examples/system_matrix_example.py:74-82 does for the reference itself."""
    f = _fixtures()
    opts, ang = list(f["sm3d_options"]), f["sm3d_angles"]
    rows, cols, vals = f["sm3d_rows"], f["sm3d_cols"], f["sm3d_vals"]
    W, H, D, Nu, Nv = (int(v) for v in opts[:5])
    x = np.random.default_rng(3).random((1, D, H, W))
    y = np.zeros(len(ang) * Nv * Nu)
    np.add.at(y, rows, vals.astype(np.float64) * x.ravel()[cols])
    want = oracle.op3d("A", x, ang, opts).ravel()
    assert rel_l2(y, want) < 3e-6


def test_2d_system_matrix_times_image_equals_projection():
    f = _fixtures()
    opts, ang = list(f["sm2d_options"]), f["sm2d_angles"]
    rows, cols, vals = f["sm2d_rows"], f["sm2d_cols"], f["sm2d_vals"]
    W, H, Nu = int(opts[0]), int(opts[1]), int(opts[2])
    x = np.random.default_rng(4).random((1, H, W))
    y = np.zeros(len(ang) * Nu)
    np.add.at(y, rows, vals.astype(np.float64) * x.ravel()[cols])
    assert rel_l2(y, oracle.op2d("A", x, ang, opts).ravel()) < 3e-6
