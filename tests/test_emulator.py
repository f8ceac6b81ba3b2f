"""CPU: the product's kernel bodies (xrecon_amd/csrc/dd_body.h), run one thread at a time
by the test-only emulator (tests/native/emu.cpp) with the product's own host plan
builder and operator routing, against the FP64 oracle.  This checks the kernel logic on
machines without a GPU; tests/test_gpu_parity.py checks the compiled gfx950 kernels."""
import numpy as np
import pytest

import oracle
from helpers import CASES2D, CASES3D, OPS, VOLUME_IN, angles_with_ties, correlate_fp64, emu_available, \
    emu_cell_box_stats, emu_cell_lines, emu_fdk3d, emu_op3d, emu_view_spans, opts2d_as_3d, opts3d, rel_l2, shapes3d, \
    uniform_angles

pytestmark = pytest.mark.skipif(not emu_available(), reason="emulator not built (run __graft_entry__.build())")



@pytest.mark.parametrize("geo", sorted(CASES3D))
@pytest.mark.parametrize("op", OPS)
def test_3d_bodies_vs_oracle(geo, op):
    opts = CASES3D[geo]
    V = 40
    ang = angles_with_ties(V)
    vol, sino = shapes3d(opts, 2, V)
    x = np.random.default_rng(1).random(vol if op in VOLUME_IN else sino).astype(np.float32)
    got = emu_op3d(op, x, ang, opts)
    want = oracle.op3d(op, x.astype(np.float64), ang, opts)
    assert rel_l2(got, want) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES2D))
@pytest.mark.parametrize("op", OPS)
def test_2d_bodies_vs_oracle(geo, op):
    o2 = CASES2D[geo]
    V = 40
    ang = angles_with_ties(V)
    rng = np.random.default_rng(2)
    if op in VOLUME_IN:
        x = rng.random((2, 17, 33)).astype(np.float32)
        got = emu_op3d(op, x[:, None], ang, opts2d_as_3d(o2), is2d=True).reshape(2, V, 45)
    else:
        x = rng.random((2, V, 45)).astype(np.float32)
        got = emu_op3d(op, x[:, :, None], ang, opts2d_as_3d(o2), is2d=True).reshape(2, 17, 33)
    want = oracle.op2d(op, x.astype(np.float64), ang, o2)
    assert rel_l2(got, want) <= 1e-5


@pytest.mark.parametrize("R,KC", [(1, 1), (2, 4), (4, 8), (8, 16), (16, 16)])
def test_tilings_give_same_results(R, KC):
    """Tile shapes change only the work split; results agree with the oracle for each."""
    opts = CASES3D["flat"]
    V = 12
    ang = angles_with_ties(V)
    vol, sino = shapes3d(opts, 1, V)
    rng = np.random.default_rng(3)
    x, y = rng.random(vol).astype(np.float32), rng.random(sino).astype(np.float32)
    assert rel_l2(emu_op3d("A", x, ang, opts, R=R), oracle.op3d("A", x.astype(np.float64), ang, opts)) <= 1e-5
    assert rel_l2(emu_op3d("B", y, ang, opts, KC=KC), oracle.op3d("B", y.astype(np.float64), ang, opts)) <= 1e-5


def test_wide_columns_take_general_path():
    """Detector columns wider than a pixel (a coarse detector) exercise the general
    multi-pixel loop of the cell gather instead of the two-pixel fast path."""
    opts = opts3d(20, 18, 6, 12, 5, 0.5, 0.5, 0.5, 4.0, 2.0, 500.0, 1000.0, 1, 0.1, 0.0, 0.0, 0.3, 0.1)
    ang = angles_with_ties(16)
    vol, sino = shapes3d(opts, 1, 16)
    rng = np.random.default_rng(4)
    x, y = rng.random(vol).astype(np.float32), rng.random(sino).astype(np.float32)
    for op, inp in (("A", x), ("BT", x), ("B", y), ("AT", y)):
        assert rel_l2(emu_op3d(op, inp, ang, opts), oracle.op3d(op, inp.astype(np.float64), ang, opts)) <= 1e-5


def test_slice_chunk_remainder_and_odd_sizes():
    """Nz not a multiple of the slice chunk, Nv not a multiple of the row tile, odd Nu."""
    opts = opts3d(21, 19, 70, 37, 13, 1.0, 1.0, 0.4, 1.0, 1.0, 500.0, 1000.0, 1)
    ang = angles_with_ties(8)
    vol, sino = shapes3d(opts, 1, 8)
    rng = np.random.default_rng(5)
    x, y = rng.random(vol).astype(np.float32), rng.random(sino).astype(np.float32)
    assert rel_l2(emu_op3d("A", x, ang, opts, R=8), oracle.op3d("A", x.astype(np.float64), ang, opts)) <= 1e-5
    assert rel_l2(emu_op3d("B", y, ang, opts, KC=16), oracle.op3d("B", y.astype(np.float64), ang, opts)) <= 1e-5


def test_volume_outside_field_of_view_gives_zero():
    """A volume shifted entirely out of the detector's field of view projects to zero."""
    opts = opts3d(8, 8, 4, 16, 4, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 1, 300.0, 0.0, 0.0)
    ang = np.array([0.0], np.float32)
    x = np.ones((1, 4, 8, 8), np.float32)
    assert np.all(emu_op3d("A", x, ang, opts) == 0.0)


@pytest.mark.parametrize("geo", sorted(CASES3D))
@pytest.mark.parametrize("variant", [1, 2])
def test_cell_gather_variants_agree(geo, variant):
    """The general cell-gather path with 32-bit byte offsets (variant 1) and with 64-bit
    entry offsets (variant 2, launched for z-prefix tables of 4 GiB and more) give the
    oracle's results on the same inputs as the per-group, per-line path selection the
    library launches."""
    opts = CASES3D[geo]
    V = 16
    ang = angles_with_ties(V)
    vol, _ = shapes3d(opts, 2, V)
    x = np.random.default_rng(6).random(vol).astype(np.float32)
    for op in ("A", "BT", "BWT"):
        want = oracle.op3d(op, x.astype(np.float64), ang, opts)
        base = emu_op3d(op, x, ang, opts)
        got = emu_op3d(op, x, ang, opts, cell_variant=variant)
        assert rel_l2(got, want) <= 1e-5
        assert rel_l2(got, base) <= 2e-6


def test_span_classes_cover_fixed_and_general_paths():
    """The benchmark's cone-flat geometry (512 x 64 detector at 1 mm, magnification 2, 1 mm
    voxels) has two-pixel views near the axes and three-pixel views near the diagonals; a
    coarse detector falls back to the general loop.  Each class agrees with the oracle."""
    cfg2 = opts3d(256, 256, 64, 512, 64, 1.0, 1.0, 1.0, 1.0, 1.0, 500.0, 1000.0, 1)
    spans = emu_view_spans(np.arange(720, dtype=np.float32) * np.float32(2 * np.pi / 720), cfg2)
    assert set(np.unique(spans)) == {2, 3}
    assert np.count_nonzero(spans == 2) > np.count_nonzero(spans == 3)
    coarse = opts3d(20, 18, 6, 12, 5, 0.5, 0.5, 0.5, 4.0, 2.0, 500.0, 1000.0, 1, 0.1, 0.0, 0.0, 0.3, 0.1)
    assert np.all(emu_view_spans(angles_with_ties(16), coarse) == 0)
    # a small geometry with both fixed classes, through every cell operator
    opts = opts3d(24, 20, 6, 30, 6, 1.0, 1.0, 1.0, 1.6, 1.0, 500.0, 1000.0, 1, 0.2, -0.1, 0.1, 0.2, 0.1)
    ang = angles_with_ties(24)
    sp = emu_view_spans(ang, opts)
    assert {2, 3} <= set(np.unique(sp))
    vol, _ = shapes3d(opts, 1, 24)
    x = np.random.default_rng(8).random(vol).astype(np.float32)
    for op in ("A", "BT", "BWT"):
        assert rel_l2(emu_op3d(op, x, ang, opts), oracle.op3d(op, x.astype(np.float64), ang, opts)) <= 1e-5


def test_span_classes_per_column_group():
    """The kernels take the span class of their 64-column group: never wider than the
    view's class, the view's class is the widest of its groups, and on diagonal views of a
    wide flat panel the central groups take the two-pixel path while the edge groups take
    three pixels.  Views with mixed groups agree with the oracle through every cell
    operator, staged and unstaged."""
    opts = opts3d(64, 60, 8, 160, 8, 1.0, 1.0, 1.0, 1.2, 1.0, 500.0, 1000.0, 1, 0.2, -0.1, 0.1, 0.2, 0.1)
    ang = angles_with_ties(48)
    spans, groups = emu_view_spans(ang, opts, groups=True)
    rank = {2: 0, 3: 1, 0: 2}
    vr = np.vectorize(rank.get)
    assert np.all(vr(groups) <= vr(spans)[:, None])
    assert np.all(vr(groups).max(axis=1) == vr(spans))
    mixed = [v for v in range(len(ang)) if len(set(groups[v])) > 1]
    assert mixed
    sub = np.ascontiguousarray(ang[mixed])
    vol, _ = shapes3d(opts, 1, len(mixed))
    x = np.random.default_rng(12).random(vol).astype(np.float32)
    for op in ("A", "BT", "BWT"):
        want = oracle.op3d(op, x.astype(np.float64), sub, opts)
        for variant in (0, 3):
            assert rel_l2(emu_op3d(op, x, sub, opts, cell_variant=variant), want) <= 1e-5, (op, variant)


# detector columns about two pixels wide: every group in the general class, its lines on the
# fixed three- and four-pixel paths
WIDE_COLUMNS = opts3d(40, 36, 6, 40, 6, 0.5, 0.5, 0.5, 2.2, 1.0, 500.0, 1000.0, 1, 0.1, -0.05, 0.0, 0.3, 0.1)


def test_wide_column_lines_take_four_pixel_path():
    ang = angles_with_ties(24)
    _, groups = emu_view_spans(ang, WIDE_COLUMNS, groups=True)
    assert np.all(groups == 0)
    x = np.random.default_rng(3).random(shapes3d(WIDE_COLUMNS, 1, 24)[0]).astype(np.float32)
    for op in ("A", "BT", "BWT"):
        want = oracle.op3d(op, x.astype(np.float64), ang, WIDE_COLUMNS)
        assert rel_l2(emu_op3d(op, x, ang, WIDE_COLUMNS), want) <= 1e-5, op
        assert rel_l2(emu_op3d(op, x, ang, WIDE_COLUMNS, cell_variant=1), want) <= 1e-5, op


@pytest.mark.parametrize("geo", sorted(CASES3D))
def test_voxel_gather_staged_and_direct_agree(geo):
    """The voxel gather with the per-view block staged in LDS (lookups outside the block
    read the table) and with every lookup from the table give the oracle's results."""
    opts = CASES3D[geo]
    V = 16
    ang = angles_with_ties(V)
    _, sino = shapes3d(opts, 2, V)
    y = np.random.default_rng(9).random(sino).astype(np.float32)
    for op in ("B", "BW", "AT"):
        want = oracle.op3d(op, y.astype(np.float64), ang, opts)
        staged = emu_op3d(op, y, ang, opts)
        direct = emu_op3d(op, y, ang, opts, voxel_variant=1)
        assert rel_l2(staged, want) <= 1e-5
        assert np.array_equal(staged, direct)


@pytest.mark.parametrize("geo", sorted(CASES3D))
def test_cell_gather_staged_equals_direct(geo):
    """The LDS-staged cell workgroups (each line's block copied into a slot whose other
    entries are NaN, then every thread's lookups served from it) give the direct threads'
    results bit for bit, every lookup inside its block; the workgroups mixing the two
    orientations (around the 45-degree ties) run directly."""
    opts = CASES3D[geo]
    V = 90
    ang = angles_with_ties(V)
    vol, _ = shapes3d(opts, 2, V)
    x = np.random.default_rng(31).random(vol).astype(np.float32)
    emu_cell_lines()
    for op in ("A", "BT", "BWT"):
        want = oracle.op3d(op, x.astype(np.float64), ang, opts)
        for R in (4, 8, 16):
            direct = emu_op3d(op, x, ang, opts, R=R)
            emu_cell_lines()
            staged = emu_op3d(op, x, ang, opts, R=R, cell_variant=3)
            lds_lines, misses, wg_staged, wg_direct = emu_cell_lines()
            assert np.array_equal(staged, direct), (op, R)
            assert misses == 0 and lds_lines > 0 and wg_staged > 0 and wg_direct > 0, (op, R)
        assert rel_l2(staged, want) <= 1e-5, op


def test_cell_gather_staged_oversize_lines():
    """Tall detector rows over thin slices: some blocks of 16-row tiles exceed a slot, and
    the plan runs those (workgroup, row tile) pairs directly; the results still equal the
    direct ones and the oracle's."""
    opts = opts3d(24, 20, 120, 48, 16, 1.0, 1.0, 0.25, 1.0, 2.0, 500.0, 1000.0, 1, 0.2, -0.1, 0.05, 0.2, 0.3)
    V = 48
    ang = uniform_angles(V)
    st = emu_cell_box_stats(ang, opts, 16, 8)
    assert 0 < st[1] < st[0] and st[2] > 0 and st[3] > 0  # some blocks over a slot; pairs of both kinds
    vol, _ = shapes3d(opts, 1, V)
    x = np.random.default_rng(37).random(vol).astype(np.float32)
    for op in ("A", "BWT"):
        direct = emu_op3d(op, x, ang, opts, R=16)
        emu_cell_lines()
        staged = emu_op3d(op, x, ang, opts, R=16, cell_variant=3)
        lds_lines, misses, _, _ = emu_cell_lines()
        assert np.array_equal(staged, direct) and misses == 0 and lds_lines > 0, op
        assert rel_l2(staged, oracle.op3d(op, x.astype(np.float64), ang, opts)) <= 1e-5, op


@pytest.mark.parametrize("case", ["cone_flat_256x256x64", "cone_arc_512x512x256"])
def test_cell_staging_blocks_of_the_workloads(case):
    """The staged blocks of the two full-size workloads, as the slot capacity comment
    states them: 256^2 x 64 cone-flat (720 views): every block of 8-row tiles fits a
    slot (616 entries at most); 512^2 x 256 cone-arc (1152 views), 8 rows: under 0.1 %
    of the blocks exceed a slot (1860 entries at most); the (workgroup, row tile) pairs
    run directly (orientations mixed, or a block over a slot) are a few per cent."""
    if case == "cone_flat_256x256x64":
        opts, V = opts3d(256, 256, 64, 512, 64, 1, 1, 1, 1.0, 1.0, 500.0, 1000.0, 1), 720
        st = emu_cell_box_stats(uniform_angles(V), opts, 8, 8)
        assert st[1] == 0 and st[5] == 616
    else:
        opts = opts3d(512, 512, 256, 736, 128, .7, .7, .7, float(np.float32(1.2 / 1085.6)), 1.0, 595.0, 1085.6, 2)
        V = 1152
        st = emu_cell_box_stats(uniform_angles(V), opts, 8, 8)
        assert 0 < st[1] / st[0] < 0.001 and st[5] == 1860
    assert st[3] / (st[2] + st[3]) < 0.05


def test_prefix_passes_over_several_chunks():
    """Columns of 70 entries (three load chunks of the prefix passes, above the register
    path's 64): the z-prefix (A, B_W^T) and row-prefix (B, A^T) tables give the oracle's
    results."""
    opts = opts3d(12, 10, 70, 20, 70, 1.0, 1.0, 0.5, 1.0, 0.5, 500.0, 1000.0, 1, 0.1, -0.1, 0.05, 0.1, 0.2)
    V = 6
    ang = angles_with_ties(V)
    vol, sino = shapes3d(opts, 1, V)
    rng = np.random.default_rng(41)
    for op in ("A", "BWT", "B", "AT"):
        x = rng.random(vol if op in VOLUME_IN else sino).astype(np.float32)
        want = oracle.op3d(op, x.astype(np.float64), ang, opts)
        assert rel_l2(emu_op3d(op, x, ang, opts), want) <= 1e-5, op


@pytest.mark.parametrize("geo", ["flat", "arc", "par"])
def test_voxel_gather_skips_views_off_the_detector(geo):
    """A tall, wide volume seen by a short, narrow detector: most (tile, slice chunk)
    pairs miss the detector on most views and the voxel gather passes over them; the
    results still equal the path that visits every view (bitwise) and the oracle."""
    import ctypes

    from helpers import _load_emu

    du = {"flat": 1.0, "arc": 1.0 / 1000.0, "par": 0.5}[geo]
    sod, sdd = (0.0, 0.0) if geo == "par" else (500.0, 1000.0)
    opts = opts3d(48, 40, 64, 24, 6, 1.0, 1.0, 1.0, du, 1.0, sod, sdd, {"flat": 1, "arc": 2, "par": 0}[geo],
                  0.3, -0.2, 0.4, 0.2, 0.1)
    V = 12
    ang = angles_with_ties(V)
    _, sino = shapes3d(opts, 1, V)
    y = np.random.default_rng(10).random(sino).astype(np.float32)
    lib = _load_emu()
    lib.emu_voxel_skips.restype = ctypes.c_longlong
    for op in ("B", "BW", "AT"):
        lib.emu_voxel_skips(1)
        staged = emu_op3d(op, y, ang, opts, KC=16)
        assert lib.emu_voxel_skips(1) > 0
        direct = emu_op3d(op, y, ang, opts, KC=16, voxel_variant=1)
        assert np.array_equal(staged, direct), op
        assert rel_l2(staged, oracle.op3d(op, y.astype(np.float64), ang, opts)) <= 1e-5
        # 16 slices as two halves, each passed over where no row of it is on the detector
        assert np.array_equal(emu_op3d(op, y, ang, opts, KC=16 | 0x100), staged), op
        assert np.array_equal(emu_op3d(op, y, ang, opts, KC=16 | 0x100, voxel_variant=1), staged), op
        # 32 slices as two halves of 16 (chunks start at other slices: equal to tolerance)
        wide = emu_op3d(op, y, ang, opts, KC=32 | 0x100)
        assert np.array_equal(wide, emu_op3d(op, y, ang, opts, KC=32 | 0x100, voxel_variant=1)), op
        assert rel_l2(wide, staged) <= 1e-6, op


def test_big_plan_cell_gather_vs_oracle():
    """A plane of 1456 x 1456 voxels (Nx Ny >= 2^21) makes the plan "big": the cell
    gathers then run the general loop with 64-bit table offsets on every line.  A, B^T and
    B_W^T of such a plan against the oracle, and against the 32-bit general path."""
    opts = opts3d(1456, 1456, 1, 40, 2, 0.2, 0.2, 1.0, 1.0, 1.0, 500.0, 1000.0, 1, 0.3, -0.2, 0.0, 0.3, 0.1)
    ang = np.array([0.1, 0.9, 2.3], dtype=np.float32)
    vol, _ = shapes3d(opts, 1, 3)
    x = np.random.default_rng(12).random(vol).astype(np.float32)
    for op in ("A", "BT", "BWT"):
        got = emu_op3d(op, x, ang, opts)
        assert rel_l2(got, oracle.op3d(op, x.astype(np.float64), ang, opts)) <= 1e-5, op
        assert rel_l2(got, emu_op3d(op, x, ang, opts, cell_variant=1)) <= 2e-6, op


def _fdk_inputs(nu, seed):
    from xrecon_amd.filters import filter_generate

    rng = np.random.default_rng(seed)
    # an asymmetric kernel, so that a reversed tap index or column order shows
    filt = (filter_generate(nu, 1.0, "shepplogan").numpy() * 0.37 + 0.05 * rng.random(2 * nu - 1)).astype(np.float32)
    weight = (0.5 + rng.random(nu)).astype(np.float32)
    return rng, filt, weight


@pytest.mark.parametrize("geo", sorted(CASES3D))
def test_fused_fdk_vs_fp64_pipeline(geo):
    """The fused FDK pass (filter stage, in-place prefix, weighted voxel gather) against
    weight -> FP64 correlation -> FP64 oracle B_W."""
    opts = CASES3D[geo]
    V = 24
    ang = angles_with_ties(V)
    _, sino = shapes3d(opts, 2, V)
    rng, filt, weight = _fdk_inputs(sino[-1], 31)
    y = rng.random(sino).astype(np.float32)
    got = emu_fdk3d(y, ang, opts, filt, weight)
    want = oracle.op3d("BW", correlate_fp64(y.astype(np.float64) * weight, filt), ang, opts)
    assert rel_l2(got, want) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES2D))
def test_fused_fbp_2d_vs_fp64_pipeline(geo):
    o2 = CASES2D[geo]
    V = 30
    ang = angles_with_ties(V)
    rng, filt, weight = _fdk_inputs(45, 32)
    y = rng.random((2, V, 45)).astype(np.float32)
    got = emu_fdk3d(y[:, :, None], ang, opts2d_as_3d(o2), filt, weight, is2d=True).reshape(2, 17, 33)
    want = oracle.op2d("BW", correlate_fp64(y.astype(np.float64) * weight, filt), ang, o2)
    assert rel_l2(got, want) <= 1e-5


@pytest.mark.parametrize("geo", sorted(CASES3D))
def test_register_prefix_bodies_match_generic(geo):
    """The prefix passes keep columns of at most 64 entries in registers on the device
    (vol_prefix_reg, sino_prefix_reg); their tables equal the generic bodies' bitwise."""
    from helpers import _load_emu, _p

    opts = CASES3D[geo]
    V = 12
    ang = angles_with_ties(V)
    vol, sino = shapes3d(opts, 2, V)
    rng = np.random.default_rng(5)
    x = rng.random(vol).astype(np.float32)
    y = rng.random(sino).astype(np.float32)
    vn, vf, gn, gf = oracle.split3d(opts)
    lib = _load_emu()
    assert lib.emu_prefix_bodies_agree(_p(vn), _p(vf), _p(gn), _p(gf), _p(ang), V, _p(x), _p(y), 2) == 1
