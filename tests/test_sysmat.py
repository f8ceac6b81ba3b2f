"""CPU: the system-matrix builder against the reference's own CPU builder.

The bar is bit-exact indices and sparsity pattern (north star), checked against the
golden triplets in tests/golden/ref_cpu_fixtures.npz and, where oracle/_ref is built,
live on random geometries of every type; values are compared exactly too (they come
out of the same FP32 expressions)."""
import numpy as np
import pytest
import torch

import oracle
import xrecon_amd
from helpers import opts2d, opts3d, rel_l2
from xrecon_amd.sysmat import system_matrix2d, system_matrix3d, triplets


def _coalesce(rows, cols, vals, shape):
    t = torch.sparse_coo_tensor(torch.stack([torch.as_tensor(rows), torch.as_tensor(cols)]),
                                torch.as_tensor(vals), shape).coalesce()
    return t.indices().numpy(), t.values().numpy()


def _check(ours: torch.Tensor, rows, cols, vals):
    ri, rv = _coalesce(rows, cols, vals, tuple(ours.shape))
    assert ours.is_coalesced()
    assert np.array_equal(ours.indices().numpy(), ri), "sparsity pattern differs"
    assert np.array_equal(ours.values().numpy(), rv), "values differ"


def test_golden_3d_and_2d():
    f = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_cpu_fixtures.npz"))
    _check(system_matrix3d(list(f["sm3d_options"]), f["sm3d_angles"]), f["sm3d_rows"], f["sm3d_cols"], f["sm3d_vals"])
    _check(system_matrix2d(list(f["sm2d_options"]), f["sm2d_angles"]), f["sm2d_rows"], f["sm2d_cols"], f["sm2d_vals"])


def _same_order(ours, rows, cols, vals):
    r, c, v = (t.numpy() for t in ours)
    assert np.array_equal(r, rows) and np.array_equal(c, cols), "triplet order or indices differ"
    assert np.array_equal(v.view(np.int32), np.asarray(vals, np.float32).view(np.int32)), "values differ"


def test_golden_triplets_in_traversal_order():
    """Before coalescing, the host builder's triplets are the reference's, in the same
    order (the GPU emission is checked against this order)."""
    f = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_cpu_fixtures.npz"))
    _same_order(triplets(list(f["sm3d_options"]), f["sm3d_angles"]), f["sm3d_rows"], f["sm3d_cols"], f["sm3d_vals"])
    _same_order(triplets(list(f["sm2d_options"]), f["sm2d_angles"], is2d=True), f["sm2d_rows"], f["sm2d_cols"],
                f["sm2d_vals"])


@pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU library not built")
@pytest.mark.parametrize("scan", [0, 1, 2])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bit_exact_vs_reference_3d(scan, seed):
    rng = np.random.default_rng(100 * scan + seed)
    W, H, D = (int(v) for v in rng.integers(4, 11, size=3))
    Nu, Nv = int(rng.integers(6, 18)), int(rng.integers(2, 7))
    du = float(rng.uniform(0.6, 1.6)) if scan != 2 else float(rng.uniform(0.6, 1.6)) / 1000.0
    o = opts3d(W, H, D, Nu, Nv, *rng.uniform(0.7, 1.3, size=3), du, float(rng.uniform(0.7, 1.4)), 500.0, 1000.0, scan,
               *rng.uniform(-0.5, 0.5, size=3), du * float(rng.uniform(-0.5, 0.5)), float(rng.uniform(-0.5, 0.5)))
    V = 9
    ang = (np.arange(V, dtype=np.float32) * np.float32(2 * np.pi / V) + np.float32(rng.uniform(0, 1))).astype(np.float32)
    ang[3] = np.float32(np.pi / 4)  # a driving-axis tie
    rows, cols, vals = oracle.ref_sm3d(ang, o)
    _check(system_matrix3d(o, ang), rows, cols, vals)
    _same_order(triplets(o, ang), rows, cols, vals)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU library not built")
@pytest.mark.parametrize("scan", [0, 1, 2])
def test_bit_exact_vs_reference_2d(scan):
    rng = np.random.default_rng(7 + scan)
    du = 0.6 if scan != 2 else 2 * np.arctan(1.2 / 2000.0)
    o = opts2d(13, 11, 21, 1.0, 1.1, du, 500.0, 1000.0, scan, 0.3, -0.2, 0.3 * du)
    ang = (np.arange(14, dtype=np.float32) * np.float32(2 * np.pi / 14) + np.float32(0.05)).astype(np.float32)
    rows, cols, vals = oracle.ref_sm2d(ang, o)
    _check(system_matrix2d(o, ang), rows, cols, vals)
    _same_order(triplets(o, ang, is2d=True), rows, cols, vals)


def test_projector_method_and_matvec():
    p = xrecon_amd.Projector3d((8, 7, 4), (12, 4), 1.0, 1.0, 500.0, 1000.0)
    ang = torch.arange(6, dtype=torch.float32) * np.float32(2 * np.pi / 6)
    M = p.system_matrix(ang)
    assert M.shape == (6 * 4 * 12, 4 * 7 * 8)
    x = np.random.default_rng(0).random((1, 4, 7, 8))
    y = torch.sparse.mm(M.double(), torch.from_numpy(x.reshape(-1, 1))).numpy().ravel()
    assert rel_l2(y, oracle.op3d("A", x, ang.numpy(), p._host_options).ravel()) < 3e-6


def test_rejects_device_tensors_and_bad_options():
    from xrecon_amd import _backend as be

    with pytest.raises(RuntimeError):
        be.native().system_matrix3d(torch.zeros(5), torch.zeros(3))
