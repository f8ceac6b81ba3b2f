"""CPU: the Python API mirror — options packing, filters, weights, FBP filtering and the
error behaviour — without GPU compute.

The filtering step of FBP/FDK (the reference runs it through torch's conv3d/conv2d,
projector3d.py:190, projector.py:201) is pinned here against a direct FP64 correlation
and the filter kernels against their closed forms and an FP64 numpy DFT."""
import math

import numpy as np
import pytest
import torch

import xrecon_amd
from xrecon_amd import filters
from xrecon_amd.geometry import ConeBeam3d, Detector3d, FanBeam2d, Detector2d, Volume2d, Volume3d


def test_projector3d_options_layout_flat():
    p = xrecon_amd.Projector3d((256, 240, 64), (512, 64), (1.0, 1.1, 0.9), (1.0, 1.2), 500.0, 1000.0,
                               pixshift=(0.5, -1.0, 2.0), binshift=(0.25, -0.5), scan_type="flat")
    o = p.options.tolist()
    # projector3d.py:98-107: W,H,D,Nu,Nv,dx,dy,dz,du(phys for flat),dv,SOD,SDD,shx*dx,shy*dy,shz*dz,
    #                        shu*du(phys),shv*dv,scan_id
    want = [256, 240, 64, 512, 64, 1.0, 1.1, 0.9, 1.0, 1.2, 500.0, 1000.0, 0.5, -1.1, 1.8, 0.25, -0.6, 1]
    assert np.allclose(o, np.float32(want), rtol=0, atol=1e-6)
    assert p.det_size == (0.5, 1.2)  # effective (virtual) pitch for weights and filters


def test_projector3d_options_layout_arc_and_parallel():
    arc = xrecon_amd.Projector3d((32, 32, 8), (64, 8), 0.7, (1.2, 1.0), 595.0, 1085.6, binshift=(0.25, 0.0),
                                 scan_type="arc")
    o = arc.options.tolist()
    assert o[17] == 2
    assert math.isclose(o[8], np.float32(1.2 / 1085.6), rel_tol=1e-6)  # arc: du/SDD radians (geometry.py:199-203)
    assert math.isclose(o[15], np.float32(0.25 * 1.2 / 1085.6), rel_tol=1e-6)
    par = xrecon_amd.Projector3d((32, 32, 8), (64, 8), 1.0, (0.9, 1.0), scan_type="parallel")
    assert par.options.tolist()[17] == 0 and math.isclose(par.options.tolist()[8], 0.9, rel_tol=1e-6)


def test_projector3d_from_value_objects_matches_legacy_constructor():
    vol = Volume3d((32, 30, 8), (1.0, 1.0, 1.0))
    geo = ConeBeam3d(Detector3d((40, 8), (1.0, 1.0)), 500.0, 1000.0, "flat")
    a = xrecon_amd.Projector3d(vol, geo)
    b = xrecon_amd.Projector3d((32, 30, 8), (40, 8), 1.0, 1.0, 500.0, 1000.0)
    assert a.options.tolist() == b.options.tolist()


def test_projector2d_options_layout():
    p = xrecon_amd.Projector2d((256, 250), 300, (1.0, 1.1), 2.0, 500.0, 1000.0, pixshift=(0.5, 1.0), binshift=0.25,
                               scan_type="flat")
    o = p.options.tolist()
    # projector.py:106-110: W,H,Nu,dx,dy,du_eff,SOD,SDD,shx*dx,shy*dy,shu_eff,scan_id
    want = [256, 250, 300, 1.0, 1.1, 1.0, 500.0, 1000.0, 0.5, 1.1, 0.25, 1]
    assert np.allclose(o, np.float32(want), atol=1e-6)
    arc = xrecon_amd.Projector2d(Volume2d(64, 1.0), FanBeam2d(Detector2d(80, 1.0), 500.0, 1000.0, "arc"))
    assert math.isclose(arc.det_size, 2 * math.atan(1.0 / 2000.0), rel_tol=1e-12)


def test_ramp_and_shepplogan_closed_forms():
    N, du = 9, 0.7
    r = filters.ramp(N, du, torch.float64).numpy()
    n = np.arange(-(N - 1), N)
    odd = -1 / (np.pi ** 2 * du * du * np.maximum(n * n, 1))
    want = np.where(n == 0, 1 / (4 * du * du), np.where(n % 2 != 0, odd, 0.0))
    assert np.allclose(r, want, rtol=1e-10, atol=0)
    s = filters.shepplogan(N, du, torch.float64).numpy()
    assert np.allclose(s, -2 / (np.pi ** 2 * du * du * (4 * n * n - 1)), rtol=1e-12)


@pytest.mark.parametrize("name,window", [("cosine", lambda w: np.cos(w / 2)),
                                         ("hamming", lambda w: 0.54 + 0.46 * np.cos(w)),
                                         ("hann", lambda w: 0.5 * (1 + np.cos(w)))])
def test_windowed_filters_vs_numpy_dft(name, window):
    N, du = 16, 0.5
    ramp = filters.ramp(N, du, torch.float32).double().numpy()
    k = np.concatenate([np.arange(0, N), np.arange(-N + 1, 0)])
    want = np.fft.ifft(np.fft.fft(ramp) * window(k * np.pi / N)).real
    got = filters.filter_generate(N, du, name).numpy()
    assert np.allclose(got, want.astype(np.float32), rtol=1e-5, atol=1e-7)


def test_filter_generate_rejects_unknown():
    with pytest.raises(ValueError):
        filters.filter_generate(8, 1.0, "butterworth")


@pytest.mark.parametrize("nu", [1, 2, 7, 64])
def test_correlate_rows_vs_direct_fp64(nu):
    rng = np.random.default_rng(nu)
    x = rng.standard_normal((2, 3, nu))
    h = rng.standard_normal(2 * nu - 1)  # asymmetric: checks correlation vs convolution
    want = np.zeros_like(x)
    for u in range(nu):
        for s in range(nu):
            want[..., u] += x[..., s] * h[s - u + nu - 1]
    got = filters.correlate_rows(torch.from_numpy(x), torch.from_numpy(h)).numpy()
    assert np.allclose(got, want, rtol=1e-10, atol=1e-10)


def test_correlate_rows_matches_conv_padding_semantics():
    """The reference filters with conv3d(sino, filt, padding=(0, 0, Nu - 1))."""
    nu = 12
    x = torch.randn(1, 1, 3, 2, nu, dtype=torch.float64)
    h = torch.randn(2 * nu - 1, dtype=torch.float64)
    want = torch.nn.functional.conv3d(x, h.view(1, 1, 1, 1, -1), padding=(0, 0, nu - 1))
    assert torch.allclose(filters.correlate_rows(x, h), want, atol=1e-10)


def test_cosine_weights_closed_form():
    p = xrecon_amd.Projector3d((16, 16, 4), (20, 4), 1.0, 1.0, 500.0, 1000.0, binshift=(0.5, 0.0))
    du, shu = p.det_size[0], p.binshift[0]
    s = (np.arange(20) - 10 + 0.5) * du + shu
    assert np.allclose(p.weight.view(-1).numpy(), 500.0 / np.sqrt(s * s + 500.0 ** 2), rtol=1e-6)
    arc = xrecon_amd.Projector3d((16, 16, 4), (20, 4), 1.0, 1.0, 500.0, 1000.0, scan_type="arc")
    s = (np.arange(20) - 10 + 0.5) * arc.det_size[0]
    assert np.allclose(arc.weight.view(-1).numpy(), np.cos(s), rtol=1e-6)


def test_parker_weights_properties_and_errors():
    p = xrecon_amd.Projector3d((16, 16, 4), (40, 4), 1.0, 1.0, 500.0, 1000.0)
    gmax = float(p._fan_angles().abs().max())
    full = torch.linspace(0, 2 * math.pi, 361)[:-1]
    assert torch.all(p.parker_weight(torch.linspace(0, 2 * math.pi, 10)) == 1)
    short = torch.linspace(0, math.pi + 2 * gmax, 200)
    w = p.parker_weight(short)
    assert w.shape == (1, 1, 200, 1, 40)
    assert float(w.min()) >= 0 and float(w.max()) <= 1
    with pytest.raises(ValueError):
        p.parker_weight(torch.linspace(0, 2.0, 50))
    assert p._mean_angle_step(full).item() == pytest.approx(2 * math.pi / 360, rel=1e-5)


@pytest.mark.parametrize("scan_type", ["flat", "arc"])
def test_parker_weights_sum_to_one_over_conjugate_rays(scan_type):
    """Parker's normalisation (Parker 1982; Silver 2000 for overscan): every line measured
    twice in a short scan of range pi + 2 delta gets weights summing to one.  The ray of
    column c at source angle beta and the ray of the mirrored column at beta + pi + 2 gamma_c
    are the same line -- checked first with the projector's own forward projection of a
    smooth blob, where the opposite sign is visibly another line -- and their weights are
    then checked to sum to one for every column, at the minimal range and with overscan."""
    p = xrecon_amd.Projector3d((64, 64, 1), (96, 1), 1.0, 1.0, 300.0, 600.0, scan_type=scan_type)
    Nu = 96
    gam = p._fan_angles(dtype=torch.float64).numpy()
    y, x = np.meshgrid(np.arange(64) - 31.5, np.arange(64) - 31.5, indexing="ij")
    blob = torch.from_numpy(np.exp(-((x - 8) ** 2 + (y + 5) ** 2) / 72.0).astype(np.float32)).view(1, 1, 1, 64, 64)
    for c in (10, 30, 70):
        b = 0.4
        ang = torch.tensor([b, b + math.pi + 2 * gam[c], b + math.pi - 2 * gam[c]], dtype=torch.float32)
        s = p.projection(blob, ang)[0, 0, :, 0, :]
        assert abs(float(s[1, Nu - 1 - c] / s[0, c]) - 1) < 2e-3
        assert abs(float(s[2, Nu - 1 - c] / s[0, c]) - 1) > 1e-2
    gmax = float(np.abs(gam).max())
    for delta in (gmax, gmax + 0.15):
        rng = math.pi + 2 * delta
        pairs = []
        live = delta - gam > 1e-9  # the end column of a minimal scan has no ramp (its rays are the scan's ends)
        for c in np.flatnonzero(live)[::5]:
            for f in (0.1, 0.35, 0.6, 0.9):
                b = f * 2 * (delta - gam[c])
                pairs.append((b, b + math.pi + 2 * gam[c], c))
        ang = sorted({0.0, rng, *(b for b, _, _ in pairs), *(b2 for _, b2, _ in pairs)})
        w = p.parker_weight(torch.tensor(ang, dtype=torch.float64))[0, 0, :, 0, :].numpy()
        pos = {a: k for k, a in enumerate(ang)}
        sums = [w[pos[b], c] + w[pos[b2], Nu - 1 - c] for b, b2, c in pairs]
        assert np.allclose(sums, 1.0, rtol=0, atol=1e-12)
        assert np.all(w[0][live] == 0) and np.all(np.abs(w[-1][live[::-1]]) < 1e-24)


def test_shape_checks_raise_value_error():
    p = xrecon_amd.Projector3d((16, 16, 4), (20, 4), 1.0, 1.0, 500.0, 1000.0)
    ang = torch.zeros(5)
    with pytest.raises(ValueError):
        p.projection(torch.zeros(1, 1, 4, 16, 15), ang)
    with pytest.raises(ValueError):
        p.backprojection(torch.zeros(1, 1, 6, 4, 20), ang)
    with pytest.raises(ValueError):
        p.backprojection(torch.zeros(1, 1, 5, 4, 20), torch.zeros(5, 1))
    with pytest.raises(ValueError):
        xrecon_amd.Projector3d((16, 16, 4), (20, 4), 1.0, 1.0, 500.0, 1000.0, scan_type="helical")
    with pytest.raises(TypeError):
        xrecon_amd.Projector3d((16, 16, 4))


def test_cpu_tensors_take_host_plans_and_devices_never_mix():
    """CPU inputs run the host executor (as the reference runs its CPU kernels for CPU
    tensors); a host plan refuses GPU-side use and a device plan refuses CPU tensors."""
    from xrecon_amd import _backend as be

    p = xrecon_amd.Projector3d((16, 16, 4), (20, 4), 1.0, 1.0, 500.0, 1000.0)
    y = p.projection(torch.ones(1, 1, 4, 16, 16), torch.arange(5, dtype=torch.float32))
    assert y.shape == (1, 1, 5, 4, 20) and not y.is_cuda and float(y.abs().sum()) > 0
    plan = be.get_plan(p._host_options, torch.arange(5, dtype=torch.float32), torch.empty(0), False)
    assert type(plan).__name__ == "HostPlan"
    p2 = xrecon_amd.Projector2d(16, 20, 1.0, 1.0, 500.0, 1000.0)
    b = p2.backprojection(torch.ones(1, 1, 5, 20), torch.arange(5, dtype=torch.float32))
    assert b.shape == (1, 1, 16, 16) and float(b.abs().sum()) > 0


def test_native_extension_loads_and_binds_reference_names():
    from xrecon_amd import _backend

    C = _backend.native()
    for name in ("forward_projection3d", "projection_transpose3d", "backward_projection3d",
                 "weighted_backward_projection3d", "backward_projection_transpose3d",
                 "weighted_backward_projection_transpose3d", "forward_projection2d", "backward_projection2d",
                 "weighted_backward_projection2d", "projection_transpose2d", "backward_projection_transpose2d",
                 "weighted_backward_projection_transpose2d", "back_projection2d", "weighted_back_projection2d"):
        assert hasattr(C, name), name
    opts = torch.tensor([16, 16, 4, 20, 4, 1, 1, 1, 1, 1, 500, 1000, 0, 0, 0, 0, 0, 1], dtype=torch.float32)
    # a CPU tensor computes on the host; shape errors raise as in the reference
    assert C.forward_projection3d(torch.zeros(1, 1, 4, 16, 16), opts, torch.zeros(3)).shape == (1, 1, 3, 4, 20)
    with pytest.raises(RuntimeError, match="shape"):
        C.forward_projection3d(torch.zeros(1, 1, 4, 16, 15), opts, torch.zeros(3))
