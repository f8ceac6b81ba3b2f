"""GPU parity: the gfx950 kernels (through the C-ABI and the torch extension) against
the FP64 oracle and the reference CPU kernels, plus full-size properties.

Tolerances: relative L2 <= 1e-5 against the FP64 oracle for every FP32 operator
(north-star bar); <= 1e-4 against the reference's own FP32 CPU arc kernels, whose
FP32 angle sums (tanf(beta + gamma)) deviate from FP64 by up to ~1e-4 at large beta
(measured in tests/test_oracle.py).
"""
import numpy as np
import pytest
import torch

import oracle
from helpers import CASES3D, OP_ID, OPS, VOLUME_IN, angles_with_ties, emu_available, emu_op3d, opts3d, rel_l2, \
    shapes3d, uniform_angles

pytestmark = pytest.mark.gpu

OPID = {"A": 0, "AT": 1, "B": 2, "BW": 3, "BT": 4, "BWT": 5}


def _native():
    from xrecon_amd import _backend as be

    return be.native()


def gpu_op(op, x_np, angles, opts, dev, is2d=False):
    C = _native()
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(np.asarray(angles, np.float32)), dev.index or 0, is2d)
    x = torch.from_numpy(np.ascontiguousarray(x_np, np.float32)).to(dev)
    nb = x.shape[0]
    x5 = x.view(nb, 1, *x.shape[1:])
    out = plan.execute(OPID[op], x5, torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(nb, *out.shape[2:])


@pytest.mark.parametrize("geo", sorted(CASES3D))
@pytest.mark.parametrize("op", OPS)
def test_small_vs_oracle(cuda, geo, op):
    opts = CASES3D[geo]
    V = 40
    ang = angles_with_ties(V)
    rng = np.random.default_rng(7)
    vol, sino = shapes3d(opts, 2, V)
    x = rng.random(vol if op in VOLUME_IN else sino).astype(np.float32)
    print(f"[gpu] small {geo} {op} shape={x.shape}", flush=True)
    got = gpu_op(op, x, ang, opts, cuda)
    want = oracle.op3d(op, x.astype(np.float64), ang, opts)
    err = rel_l2(got, want)
    assert err <= 1e-5, f"{geo} {op}: relL2 vs FP64 oracle {err:.3e}"
    if emu_available():
        assert rel_l2(got, emu_op3d(op, x, ang, opts)) <= 1e-5


@pytest.mark.parametrize("geo", ["flat", "arc"])
@pytest.mark.parametrize("op", ["A", "B", "BW", "AT", "BT"])
def test_small_vs_reference_cpu(cuda, geo, op):
    if not oracle.ref_available():
        pytest.skip("reference CPU library not built")
    opts = CASES3D[geo]
    V = 24
    ang = uniform_angles(V)
    rng = np.random.default_rng(8)
    vol, sino = shapes3d(opts, 1, V)
    x = rng.random(vol if op in VOLUME_IN else sino).astype(np.float32)
    got = gpu_op(op, x, ang, opts, cuda)
    ref = oracle.ref_op3d(op, x, ang, opts)
    tol = 2e-5 if geo == "flat" else 2e-4
    assert rel_l2(got, ref) <= tol


@pytest.mark.parametrize("tile_rows,voxel_chunk", [(1, 1), (2, 4), (4, 8), (8, 16), (16, 16), (16, 16 | 0x100),
                                                   (16, 32 | 0x100)])
def test_tilings_agree(cuda, tile_rows, voxel_chunk):
    C = _native()
    opts = CASES3D["flat"]
    V = 20
    ang = angles_with_ties(V)
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    rng = np.random.default_rng(3)
    vol, sino = shapes3d(opts, 1, V)
    x = torch.from_numpy(rng.random(vol).astype(np.float32)).to(cuda).unsqueeze(1)
    y = torch.from_numpy(rng.random(sino).astype(np.float32)).to(cuda).unsqueeze(1)
    s = torch.cuda.current_stream().cuda_stream
    outs = {}
    for op, inp in ((0, x), (2, y)):
        out = torch.empty(plan.execute(op, inp, s).shape, device=cuda)
        ws = torch.empty(plan.workspace_bytes(op, 1), dtype=torch.uint8, device=cuda)
        plan.execute_into(op, inp, out, ws, s, tile_rows, voxel_chunk, 0)
        outs[op] = out.cpu().numpy()[:, 0]
    want_a = oracle.op3d("A", x.cpu().numpy()[:, 0].astype(np.float64), ang, opts)
    want_b = oracle.op3d("B", y.cpu().numpy()[:, 0].astype(np.float64), ang, opts)
    assert rel_l2(outs[0], want_a) <= 1e-5
    assert rel_l2(outs[2], want_b) <= 1e-5


def test_cell_classes_mixed_groups(cuda):
    """A wide flat panel whose diagonal views mix two- and three-pixel column groups: the
    cell gathers launch one kernel per class and every operator matches the oracle."""
    from helpers import opts3d

    C = _native()
    opts = opts3d(64, 60, 8, 160, 8, 1.0, 1.0, 1.0, 1.2, 1.0, 500.0, 1000.0, 1, 0.2, -0.1, 0.1, 0.2, 0.1)
    V = 48
    ang = angles_with_ties(V)
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    n2, n3, n0 = plan.cell_classes()
    assert n2 + n3 + n0 == V * 3 and n2 > 0 and n3 > 0
    rng = np.random.default_rng(17)
    vol, _ = shapes3d(opts, 2, V)
    x = rng.random(vol).astype(np.float32)
    for op in ("A", "BT", "BWT"):
        print(f"[gpu] mixed cell classes {op}", flush=True)
        got = gpu_op(op, x, ang, opts, cuda)
        assert rel_l2(got, oracle.op3d(op, x.astype(np.float64), ang, opts)) <= 1e-5, op


def test_cell_tilings_bitwise_equal(cuda):
    """Rows per thread change only the work split of the cell gather: every row is
    computed from its absolute index with the same explicit operations, so all tilings
    (and the classes' different row counts) give bitwise-equal sinograms."""
    from helpers import opts3d

    C = _native()
    opts = opts3d(64, 60, 8, 160, 8, 1.0, 1.0, 1.0, 1.2, 1.0, 500.0, 1000.0, 1, 0.2, -0.1, 0.1, 0.2, 0.1)
    V = 48
    ang = angles_with_ties(V)
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    x = torch.from_numpy(np.random.default_rng(23).random(shapes3d(opts, 1, V)[0]).astype(np.float32)).to(cuda)
    x = x.unsqueeze(1)
    s = torch.cuda.current_stream().cuda_stream
    ws = torch.empty(plan.workspace_bytes(0, 1), dtype=torch.uint8, device=cuda)
    outs = []
    for R in (1, 2, 4, 8, 16):
        out = torch.empty((1, 1, V, 8, 160), device=cuda)
        plan.execute_into(0, x, out, ws, s, R, 0, 0)
        outs.append(out)
    for R, o in zip((2, 4, 8, 16), outs[1:]):
        assert torch.equal(o, outs[0]), f"tile_rows {R} differs from tile_rows 1"


XR_CELL_DIRECT = 0x100
XR_CELL_LDS = 0x200

# Staging geometries: three column groups per view (two- and three-pixel classes on the
# flat panel), 720 views with the 45-degree ties, 4 batch items, so that every tiling
# launches per class (the staged kernels) rather than in one mixed grid.
STAGED = {
    "flat": opts3d(64, 60, 16, 160, 24, 1.0, 1.0, 1.0, 1.2, 1.0, 500.0, 1000.0, 1, 0.2, -0.1, 0.1, 0.2, 0.1),
    "arc": opts3d(64, 60, 16, 160, 24, 1.0, 1.0, 1.0, 1.2 / 1000.0, 1.0, 500.0, 1000.0, 2, 0.2, -0.1, 0.1,
                  0.2 / 1000.0, 0.1),
    "par": opts3d(64, 60, 16, 160, 24, 1.0, 1.0, 1.0, 0.9, 1.0, 0.0, 0.0, 0, 0.2, -0.1, 0.1, 0.2, 0.1),
}


@pytest.mark.parametrize("geo", sorted(STAGED))
def test_cell_staged_equals_direct(cuda, geo):
    """The cell gathers that stage each line's block of the z-prefix table in LDS read the
    same entries as the direct ones (a line whose block exceeds a slot is read from the
    table), so every operator and tiling is bitwise equal to tile_rows | XR_CELL_DIRECT.
    The default for these small tables is the direct path."""
    C = _native()
    opts = STAGED[geo]
    V, nb = 720, 4
    ang = angles_with_ties(V)
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    n2, n3, n0 = plan.cell_classes()
    assert n2 + n3 > 0
    vol, sino = shapes3d(opts, nb, V)
    rng = np.random.default_rng(29)
    s = torch.cuda.current_stream().cuda_stream
    x = torch.from_numpy(rng.random(vol).astype(np.float32)).to(cuda).unsqueeze(1)
    for op in ("A", "BT", "BWT"):  # the cell-gather operators: volume in, sinogram out
        ws = torch.empty(max(1, plan.workspace_bytes(OPID[op], nb)), dtype=torch.uint8, device=cuda)
        for R in (4, 8, 16):
            print(f"[gpu] staged vs direct {geo} {op} rows {R}", flush=True)
            outs = []
            for flag in (XR_CELL_LDS, XR_CELL_DIRECT, 0):
                out = torch.empty((nb, 1, *sino[1:]), device=cuda)
                plan.execute_into(OPID[op], x, out, ws, s, R | flag, 0, 0)
                outs.append(out)
            torch.cuda.synchronize()
            assert torch.equal(outs[0], outs[1]), f"{geo} {op} rows {R}: staged differs from direct"
            assert torch.equal(outs[2], outs[1])
            assert torch.isfinite(outs[0]).all()


def test_cell_four_pixel_lines(cuda):
    """Detector columns about two pixels wide: general-class groups whose lines take the
    fixed three- and four-pixel paths, against the oracle."""
    from helpers import opts3d

    opts = opts3d(40, 36, 6, 40, 6, 0.5, 0.5, 0.5, 2.2, 1.0, 500.0, 1000.0, 1, 0.1, -0.05, 0.0, 0.3, 0.1)
    V = 24
    ang = angles_with_ties(V)
    x = np.random.default_rng(3).random(shapes3d(opts, 2, V)[0]).astype(np.float32)
    for op in ("A", "BT", "BWT"):
        print(f"[gpu] four-pixel lines {op}", flush=True)
        got = gpu_op(op, x, ang, opts, cuda)
        assert rel_l2(got, oracle.op3d(op, x.astype(np.float64), ang, opts)) <= 1e-5, op


def test_cell_general_class_split_launch(cuda):
    """A coarse detector (every column group in the general class) with enough views that
    the cell gathers launch per class rather than in one mixed grid."""
    from helpers import opts3d

    C = _native()
    opts = opts3d(20, 18, 6, 12, 5, 0.5, 0.5, 0.5, 4.0, 2.0, 500.0, 1000.0, 1, 0.1, 0.0, 0.0, 0.3, 0.1)
    V = 4200
    ang = uniform_angles(V)
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    assert plan.cell_classes() == [0, 0, V]
    x = np.random.default_rng(19).random(shapes3d(opts, 1, V)[0]).astype(np.float32)
    for op in ("A", "BWT"):
        print(f"[gpu] general-class split launch {op}", flush=True)
        got = gpu_op(op, x, ang, opts, cuda)
        assert rel_l2(got, oracle.op3d(op, x.astype(np.float64), ang, opts)) <= 1e-5, op


# ------------------------------------------------------------------ full size (cfg 2)

CFG2 = [256, 256, 64, 512, 64, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 0, 0, 0, 0, 0, 1]


@pytest.fixture(scope="module")
def cfg2(cuda):
    from xrecon_amd.phantom import random_ellipsoids

    V = 720
    ang = uniform_angles(V)
    ph = random_ellipsoids((256, 256, 64), (1.0, 1.0, 1.0), count=16, semi_axes=(5.0, 80.0), seed=1002)
    C = _native()
    plan = C.Plan([float(v) for v in CFG2], torch.from_numpy(ang), 0, False)
    x = torch.from_numpy(ph).to(cuda).view(1, 1, 64, 256, 256)
    return plan, ang, ph, x


@pytest.mark.parametrize("rows", [8, 16])
def test_cfg2_staged_equals_direct(cfg2, cuda, rows):
    """Full-size projection: the LDS-staged cell gather equals the direct one bitwise."""
    plan, ang, ph, x = cfg2
    print(f"[gpu] cfg2 A staged vs direct, rows {rows}", flush=True)
    s = torch.cuda.current_stream().cuda_stream
    ws = torch.empty(plan.workspace_bytes(0, 1), dtype=torch.uint8, device=cuda)
    outs = []
    for flag in (XR_CELL_LDS, XR_CELL_DIRECT):
        out = torch.empty((1, 1, 720, 64, 512), device=cuda)
        plan.execute_into(0, x, out, ws, s, rows | flag, 0, 0)
        outs.append(out)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])


def test_cfg2_projection_sampled_views(cfg2, cuda):
    plan, ang, ph, x = cfg2
    print("[gpu] cfg2 A sampled views", flush=True)
    y = plan.execute(0, x, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    views = np.array([0, 1, 90, 179, 180, 270, 359, 450, 541, 719])  # includes the 45/135/225/315-degree ties
    want = oracle.op3d("A", ph[None].astype(np.float64), np.ascontiguousarray(ang[views]), CFG2)
    got = y.cpu().numpy()[0, 0][views]
    assert rel_l2(got, want[0]) <= 1e-5


def test_cfg2_backprojection_subset(cfg2, cuda):
    plan, ang, ph, x = cfg2
    views = np.array([0, 45, 90, 135, 180, 225, 270, 315, 360, 600])
    sub = np.ascontiguousarray(ang[views])
    C = _native()
    psub = C.Plan([float(v) for v in CFG2], torch.from_numpy(sub), 0, False)
    rng = np.random.default_rng(11)
    y = rng.random((1, len(views), 64, 512)).astype(np.float32)
    s = torch.cuda.current_stream().cuda_stream
    for op in ("B", "BW"):
        print(f"[gpu] cfg2 {op} on {len(views)} views", flush=True)
        got = psub.execute(OPID[op], torch.from_numpy(y).to(cuda).unsqueeze(1), s).cpu().numpy()[:, 0]
        want = oracle.op3d(op, y.astype(np.float64), sub, CFG2)
        assert rel_l2(got, want) <= 1e-5, op


@pytest.mark.parametrize("pair", [("A", "AT"), ("B", "BT"), ("BW", "BWT")])
def test_cfg2_adjointness(cfg2, cuda, pair):
    """<A x, y> = <x, A^T y> at full size (dot products in FP64 of FP32 results)."""
    plan, ang, ph, x = cfg2
    fwd, adj = pair
    s = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device=cuda).manual_seed(5)
    vol = torch.rand((1, 1, 64, 256, 256), device=cuda, generator=g)
    sino = torch.rand((1, 1, 720, 64, 512), device=cuda, generator=g)
    if fwd == "A":
        ax = plan.execute(OPID[fwd], vol, s)
        aty = plan.execute(OPID[adj], sino, s)
        lhs = (ax.double() * sino.double()).sum().item()
        rhs = (vol.double() * aty.double()).sum().item()
    else:  # B maps sinogram -> volume
        by = plan.execute(OPID[fwd], sino, s)
        btx = plan.execute(OPID[adj], vol, s)
        lhs = (by.double() * vol.double()).sum().item()
        rhs = (sino.double() * btx.double()).sum().item()
    assert abs(lhs - rhs) / abs(rhs) <= 1e-5, (lhs, rhs)


def test_cfg2_batch_items_in_split_launches(cfg2, cuda):
    """cfg2 with two batch items: the per-class cell launches and the voxel gather decode
    the batch item from the launch grid; each item equals its own single-item run."""
    plan, ang, ph, x = cfg2
    s = torch.cuda.current_stream().cuda_stream
    x2 = torch.cat([x, torch.rand_like(x)], dim=0)
    print("[gpu] cfg2 A and B with two batch items", flush=True)
    y2 = plan.execute(0, x2, s)
    for i in range(2):
        assert torch.equal(y2[i:i + 1], plan.execute(0, x2[i:i + 1].contiguous(), s)), f"A item {i}"
    b2 = plan.execute(2, y2, s)
    for i in range(2):
        assert torch.equal(b2[i:i + 1], plan.execute(2, y2[i:i + 1].contiguous(), s)), f"B item {i}"


def test_cfg2_deterministic_and_linear(cfg2, cuda):
    plan, ang, ph, x = cfg2
    s = torch.cuda.current_stream().cuda_stream
    y1 = plan.execute(0, x, s)
    y2 = plan.execute(0, x, s)
    assert torch.equal(y1, y2), "projection is not bitwise reproducible"
    b1 = plan.execute(2, y1, s)
    b2 = plan.execute(2, y1, s)
    assert torch.equal(b1, b2), "backprojection is not bitwise reproducible"
    z = torch.rand_like(x)
    lin = plan.execute(0, 2.0 * x + z, s)
    ref = 2.0 * y1 + plan.execute(0, z, s)
    err = ((lin - ref).double().norm() / ref.double().norm()).item()
    assert err <= 1e-6


def test_batch_independence(cuda):
    C = _native()
    opts = CASES3D["flat"]
    V = 16
    ang = uniform_angles(V)
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    vol, _ = shapes3d(opts, 3, V)
    x = torch.rand(vol, device=cuda).view(3, 1, *vol[1:])
    s = torch.cuda.current_stream().cuda_stream
    yb = plan.execute(0, x, s)
    for i in range(3):
        yi = plan.execute(0, x[i:i + 1].contiguous(), s)
        assert torch.equal(yb[i:i + 1], yi)
    # B = 2, C = 2 folds into four instances
    x4 = torch.rand((2, 2) + tuple(vol[1:]), device=cuda)
    y4 = plan.execute(0, x4, s)
    assert y4.shape == (2, 2, V, int(opts[4]), int(opts[3]))
    assert torch.equal(y4[1, 0], plan.execute(0, x4[1:2, 0:1].contiguous(), s)[0, 0])


@pytest.mark.parametrize("case", ["plane", "table"])
def test_big_plans_vs_oracle(cuda, case):
    """Plans whose z-prefix addressing needs 64 bits ("big": a plane of Nx Ny >= 2^21
    voxels, or a table of one batch item of 4 GiB and more) run the cell gathers'
    general loop with 64-bit offsets; A, B^T and B_W^T against the oracle, and the voxel
    gathers of the same plans.  (The CPU emulator checks the same path first,
    tests/test_emulator.py.)"""
    if case == "plane":
        opts = opts3d(1456, 1456, 2, 64, 4, 0.2, 0.2, 1.0, 1.0, 1.0, 500.0, 1000.0, 1, 0.3, -0.2, 0.0, 0.3, 0.1)
    else:  # (Nz + 1) * Nx * Ny * 8 = 4.6 GB
        opts = opts3d(800, 800, 900, 48, 24, 0.3, 0.3, 0.3, 1.0, 1.0, 500.0, 1000.0, 1, 0.0, 0.0, 0.0, 0.0, 0.0)
    ang = np.array([0.1, 0.9, 2.3], dtype=np.float32) if case == "plane" else np.array([0.3, 1.2], dtype=np.float32)
    ops = ("A", "BT", "BWT", "B", "AT") if case == "plane" else ("A", "B")
    from xrecon_amd import _backend as be

    plan = be.native().Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    s = torch.cuda.current_stream().cuda_stream
    vol, sino = shapes3d(opts, 1, len(ang))
    rng = np.random.default_rng(13)
    x = rng.random(vol, dtype=np.float32)
    y = rng.random(sino).astype(np.float32)
    xd = torch.from_numpy(x).to(cuda).unsqueeze(1)
    yd = torch.from_numpy(y).to(cuda).unsqueeze(1)
    for op in ops:
        print(f"[gpu] big plan ({case}) {op}", flush=True)
        inp = (xd, x) if op in ("A", "BT", "BWT") else (yd, y)
        got = plan.execute(OP_ID[op], inp[0], s).cpu().numpy()[:, 0]
        want = oracle.op3d(op, inp[1].astype(np.float64), ang, opts)
        err = rel_l2(got, want)
        print(f"[gpu] big plan ({case}) {op}: rel L2 {err:.2e}", flush=True)
        assert err <= 1e-5, op


def test_invalid_tile_shapes_are_refused(cuda):
    """Tile shapes outside the supported set fail with the library's message before any
    kernel runs, for a plan and for a batch (which takes no paired or unknown chunks)."""
    from xrecon_amd.batched import BatchedProjector3d
    import xrecon_amd

    C = _native()
    opts = CASES3D["flat"]
    V = 8
    ang = angles_with_ties(V)
    plan = C.Plan([float(v) for v in opts], torch.from_numpy(ang), 0, False)
    vol, sino = shapes3d(opts, 1, V)
    y = torch.rand(sino, device=cuda).unsqueeze(1)
    x = torch.rand(vol, device=cuda).unsqueeze(1)
    s = torch.cuda.current_stream().cuda_stream
    out_v = torch.empty(plan.execute(2, y, s).shape, device=cuda)
    out_s = torch.empty(plan.execute(0, x, s).shape, device=cuda)
    ws = torch.empty(max(plan.workspace_bytes(2, 1), plan.workspace_bytes(0, 1)), dtype=torch.uint8, device=cuda)
    for chunk in (2, 3, 32, 64, 8 | 0x100, 16 | 0x200, 0x100):
        with pytest.raises(RuntimeError, match="voxel_chunk must be"):
            plan.execute_into(2, y, out_v, ws, s, 0, chunk, 0)
    for rows in (3, 32, 3 | XR_CELL_DIRECT, 8 | 0x400, 8 | XR_CELL_DIRECT | XR_CELL_LDS):
        with pytest.raises(RuntimeError, match="tile_rows must be"):
            plan.execute_into(0, x, out_s, ws, s, rows, 0, 0)
    proj = xrecon_amd.Projector3d((24, 20, 8), (40, 8), 1.0, 1.0, 500.0, 1000.0)
    bp = BatchedProjector3d([proj], [torch.from_numpy(uniform_angles(9))], device=cuda)
    with pytest.raises(RuntimeError, match="voxel_chunk must be"):
        bp.run(2, [torch.rand(bp.sino_shape(0), device=cuda)], tiles=(0, 5))
