"""Seeded analytic phantoms for tests and benchmarks: sums of random ellipses (2D) or
ellipsoids (3D) evaluated at pixel/voxel centres, added together as the reference's
``xrecon/utils/phantom.py:43-50`` does for its Shepp-Logan tables.

Parameters follow the benchmark workloads: centres uniform within +-0.7 of the half
field of view, semi-axes uniform in a range (mm), orientation uniform, intensities
uniform in a range.  Seeds are ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import numpy as np


def _grid(n: int, d: float, flip: bool = False) -> np.ndarray:
    c = (np.arange(n) - n / 2 + 0.5) * d
    return c[::-1].copy() if flip else c


def random_ellipses(size, spacing=(1.0, 1.0), count=10, semi_axes=(5.0, 60.0), intensity=(0.1, 1.0),
                    seed=1001) -> np.ndarray:
    """float32 image [H, W] (row 0 at +y, as the projectors index it)."""
    W, H = size
    dx, dy = spacing
    rng = np.random.default_rng(seed)
    x = _grid(W, dx)[None, :]
    y = _grid(H, dy, flip=True)[:, None]
    half = 0.5 * min(W * dx, H * dy)
    img = np.zeros((H, W), dtype=np.float64)
    for _ in range(count):
        cx, cy = rng.uniform(-0.7, 0.7, size=2) * half
        a, b = rng.uniform(*semi_axes, size=2)
        phi = rng.uniform(0.0, np.pi)
        val = rng.uniform(*intensity)
        c, s = np.cos(phi), np.sin(phi)
        u = (x - cx) * c + (y - cy) * s
        v = -(x - cx) * s + (y - cy) * c
        img += val * ((u / a) ** 2 + (v / b) ** 2 <= 1.0)
    return img.astype(np.float32)


def random_ellipsoids(size, spacing=(1.0, 1.0, 1.0), count=16, semi_axes=(5.0, 80.0), intensity=(0.1, 1.0),
                      seed=1002, device=None):
    """float32 volume [D, H, W] (x fastest, row 0 at +y).  With `device` (a torch device)
    the same ellipsoids are evaluated there and a torch tensor is returned: large volumes
    (512^2 x 256) take seconds instead of minutes."""
    if device is not None:
        return _random_ellipsoids_torch(size, spacing, count, semi_axes, intensity, seed, device)
    W, H, D = size
    dx, dy, dz = spacing
    rng = np.random.default_rng(seed)
    x = _grid(W, dx)[None, None, :]
    y = _grid(H, dy, flip=True)[None, :, None]
    z = _grid(D, dz)[:, None, None]
    half = np.array([W * dx, H * dy, D * dz]) * 0.5
    vol = np.zeros((D, H, W), dtype=np.float32)
    for _ in range(count):
        cx, cy, cz = rng.uniform(-0.7, 0.7, size=3) * half
        ax = rng.uniform(*semi_axes, size=3)
        # random rotation: z-axis angle, then tilt about x
        phi, theta = rng.uniform(0.0, np.pi), rng.uniform(0.0, np.pi)
        val = np.float32(rng.uniform(*intensity))
        cp, sp, ct, st = np.cos(phi), np.sin(phi), np.cos(theta), np.sin(theta)
        X, Y, Z = x - cx, y - cy, z - cz
        u = X * cp + Y * sp
        v0 = -X * sp + Y * cp
        v = v0 * ct + Z * st
        w = -v0 * st + Z * ct
        vol += val * (((u / ax[0]) ** 2 + (v / ax[1]) ** 2 + (w / ax[2]) ** 2) <= 1.0).astype(np.float32)
    return vol


def _random_ellipsoids_torch(size, spacing, count, semi_axes, intensity, seed, device):
    import torch

    W, H, D = size
    dx, dy, dz = spacing
    rng = np.random.default_rng(seed)
    f64 = dict(dtype=torch.float64, device=device)
    x = torch.from_numpy(_grid(W, dx)).to(**f64).view(1, 1, W)
    y = torch.from_numpy(_grid(H, dy, flip=True)).to(**f64).view(1, H, 1)
    z = torch.from_numpy(_grid(D, dz)).to(**f64).view(D, 1, 1)
    half = np.array([W * dx, H * dy, D * dz]) * 0.5
    vol = torch.zeros((D, H, W), dtype=torch.float32, device=device)
    for _ in range(count):
        cx, cy, cz = rng.uniform(-0.7, 0.7, size=3) * half
        ax = rng.uniform(*semi_axes, size=3)
        phi, theta = rng.uniform(0.0, np.pi), rng.uniform(0.0, np.pi)
        val = float(np.float32(rng.uniform(*intensity)))
        cp, sp, ct, st = np.cos(phi), np.sin(phi), np.cos(theta), np.sin(theta)
        X, Y, Z = x - cx, y - cy, z - cz
        u = X * cp + Y * sp
        v0 = -X * sp + Y * cp
        v = v0 * ct + Z * st
        w = -v0 * st + Z * ct
        inside = (u / ax[0]) ** 2 + (v / ax[1]) ** 2 + (w / ax[2]) ** 2 <= 1.0
        vol += val * inside.to(torch.float32)
    return vol
