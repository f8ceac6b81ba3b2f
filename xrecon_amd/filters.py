"""Reconstruction filters (the reference's ``xrecon/projector/filters.py``), as spatial
kernels of length ``2 * num_det - 1`` centred at index ``num_det - 1``.

* ``ramp`` (Ram-Lak, ``filters.py:20-32``): h[0] = 1/(4 du^2), h[n odd] = -1/(pi^2 du^2 n^2),
  h[n even != 0] = 0.
* ``shepplogan`` (``filters.py:35-41``): h[n] = -2 / (pi^2 du^2 (4 n^2 - 1)).
* ``cosine`` / ``hamming`` / ``hann`` (``filters.py:44-71``): the ramp kernel multiplied in
  the DFT domain (length 2N-1, FP64) by the window w(omega), omega_k = pi k / N for
  k = 0..N-1 followed by pi (k - N + 1) / N for the negative half; real part of the
  inverse DFT, cast to float32.
"""
from __future__ import annotations

import math
from typing import Callable, Dict

import torch

FILTERS: Dict[str, Callable] = {}


def register_filter(name: str):
    def deco(fn):
        FILTERS[name] = fn
        return fn
    return deco


def filter_generate(num_det: int, det_size: float, filter_type: str = "ramp", dtype=None) -> torch.Tensor:
    if filter_type not in FILTERS:
        raise ValueError(f"Filter type {filter_type} not supported. Available: {list(FILTERS.keys())}")
    return FILTERS[filter_type](num_det, det_size, dtype)


def _taps(num_det: int) -> torch.Tensor:
    return torch.arange(-(num_det - 1), num_det, dtype=torch.float64)


@register_filter("ramp")
def ramp(num_det: int, det_size: float, dtype=None) -> torch.Tensor:
    """Ram-Lak ramp filter."""
    n = _taps(num_det)
    h = torch.zeros_like(n)
    odd = (n.abs() % 2) == 1
    # The reference adds 1e-12 to n^2 in the odd taps (filters.py:29); n^2 >= 1 there, so
    # the result is the same to float32 precision.
    h[odd] = -1.0 / (math.pi ** 2 * det_size ** 2 * (n[odd] ** 2 + 1e-12))
    h[num_det - 1] = 1.0 / (4.0 * det_size ** 2)
    return h.to(dtype or torch.float32)


@register_filter("shepplogan")
def shepplogan(num_det: int, det_size: float, dtype=None) -> torch.Tensor:
    """Shepp-Logan filter."""
    n = _taps(num_det)
    h = -2.0 / (math.pi ** 2 * det_size ** 2 * (4.0 * n ** 2 - 1.0))
    return h.to(dtype or torch.float32)


def _windowed(num_det: int, det_size: float, dtype, window: Callable[[torch.Tensor], torch.Tensor]) -> torch.Tensor:
    base = ramp(num_det, det_size, dtype).double()
    k = torch.cat((torch.arange(0, num_det), torch.arange(-num_det + 1, 0))).double()
    omega = k * math.pi / num_det
    spec = torch.fft.fft(base) * window(omega)
    return torch.fft.ifft(spec).real.float()


@register_filter("cosine")
def cosine(num_det: int, det_size: float, dtype=None) -> torch.Tensor:
    return _windowed(num_det, det_size, dtype, lambda w: torch.cos(w / 2))


@register_filter("hamming")
def hamming(num_det: int, det_size: float, dtype=None) -> torch.Tensor:
    return _windowed(num_det, det_size, dtype, lambda w: 0.54 + 0.46 * torch.cos(w))


@register_filter("hann")
def hann(num_det: int, det_size: float, dtype=None) -> torch.Tensor:
    return _windowed(num_det, det_size, dtype, lambda w: 0.5 * (1.0 + torch.cos(w)))


def correlate_rows(sino: torch.Tensor, filt: torch.Tensor) -> torch.Tensor:
    """Zero-padded correlation of every detector row (last axis, length Nu) with a
    ``2 Nu - 1`` tap kernel centred at ``Nu - 1``:

        out[u] = sum_s sino[s] * filt[s - u + Nu - 1]

    which is what the reference's ``conv3d(sino, filt, padding=(0, 0, Nu - 1))``
    (``projector3d.py:190``) and 2D ``conv2d`` (``projector.py:201``) compute.  Done as one
    real FFT product of size >= 2 Nu - 1 (O(Nu log Nu) per row instead of the direct
    2 Nu - 1 taps).  Differentiable in both arguments.
    """
    nu = sino.shape[-1]
    h = filt.reshape(-1)
    if h.numel() != 2 * nu - 1:
        raise ValueError(f"filter must have 2*Nu-1 = {2 * nu - 1} taps, got {h.numel()}")
    m = 1 << (2 * nu - 1).bit_length()
    # convolution kernel g[n] = filt[Nu - 1 - n] for n in [-(Nu-1), Nu-1], stored circularly
    g = torch.zeros(m, dtype=sino.dtype, device=sino.device)
    hr = h.to(device=sino.device, dtype=sino.dtype).flip(0)  # hr[j] = filt[2Nu-2-j], n = j - (Nu-1)
    g[:nu] = hr[nu - 1:]
    g[m - (nu - 1):] = hr[:nu - 1]
    spec = torch.fft.rfft(sino, n=m) * torch.fft.rfft(g, n=m)
    return torch.fft.irfft(spec, n=m)[..., :nu]


def fused_fdk_ok(sino: torch.Tensor, filt: torch.Tensor, weight: torch.Tensor) -> bool:
    """Whether FBP/FDK may take the fused native pass: float32 GPU data, a detector of at
    most 1024 columns, and no gradient wanted through the sinogram, filter or weights
    (the composite path is the differentiable one)."""
    if not (sino.is_cuda and sino.dtype == torch.float32 and sino.shape[-1] <= 1024):
        return False
    if torch.is_grad_enabled() and (sino.requires_grad or filt.requires_grad or weight.requires_grad):
        return False
    return True
