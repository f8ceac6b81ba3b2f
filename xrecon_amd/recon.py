"""Differentiable reconstruction on top of the projector operators.

* ``tv`` — smoothed isotropic total variation of a [B, C, D, H, W] (or [B, C, H, W])
  volume, differentiable.
* ``simulate_transmission`` — noisy line integrals from a phantom:
  ``y = -ln(max(Poisson(I0 * exp(-A x)), 1) / I0)``.  The clamp at one count is this
  module's recorded choice: at I0 = 1e5 and line integrals far above ln(I0) most counts
  are zero, and the logarithm needs a floor.
* ``reconstruct`` — model-based reconstruction ``min 0.5 ||A x - y||^2 + lam TV(x)`` with
  Adam, gradients through the hand-written adjoint (A^T) kernels.
* ``UnrolledGradientDescent`` — K unrolled steps ``x <- x - a_k A^T (A x - y)`` with
  learnable step sizes, each step wrapped in activation checkpointing so only the
  iterates are kept and the per-step sinograms are recomputed in the backward pass.
"""
from __future__ import annotations

from typing import Optional

import torch
import torch.nn as nn
from torch.utils.checkpoint import checkpoint


def tv(x: torch.Tensor, eps: float = 1e-8) -> torch.Tensor:
    """sum over voxels of sqrt(|forward differences|^2 + eps), replicate boundary."""
    dims = range(2, x.dim())
    sq = torch.zeros_like(x)
    for d in dims:
        diff = torch.diff(x, dim=d, append=x.narrow(d, x.shape[d] - 1, 1))
        sq = sq + diff * diff
    return torch.sqrt(sq + eps).sum()


def simulate_transmission(projector, x_true: torch.Tensor, angles: torch.Tensor, I0: float = 1e5,
                          seed: int = 0) -> torch.Tensor:
    with torch.no_grad():
        p = projector.projection(x_true, angles)
        g = torch.Generator(device=p.device).manual_seed(seed)
        counts = torch.poisson(I0 * torch.exp(-p.double()), generator=g)
        return (-torch.log(torch.clamp(counts, min=1.0) / I0)).float()


def data_loss(projector, x: torch.Tensor, y: torch.Tensor, angles: torch.Tensor) -> torch.Tensor:
    r = projector.projection(x, angles) - y
    return 0.5 * (r * r).sum()


def reconstruct(projector, y: torch.Tensor, angles: torch.Tensor, steps: int = 50, lr: float = 1e-2,
                tv_weight: float = 1e-3, x0: Optional[torch.Tensor] = None, callback=None) -> torch.Tensor:
    """Adam on 0.5 ||A x - y||^2 + tv_weight * TV(x), x fp32 with requires_grad."""
    W, H, D = projector.image_size
    if x0 is None:
        x0 = torch.zeros((y.shape[0], y.shape[1], D, H, W), device=y.device)
    x = x0.clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    for k in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = data_loss(projector, x, y, angles) + tv_weight * tv(x)
        loss.backward()
        opt.step()
        if callback is not None:
            callback(k, loss.detach())
    return x.detach()


class UnrolledGradientDescent(nn.Module):
    """x_{k+1} = x_k - a_k A^T (A x_k - y), k = 0..K-1, a_k learnable (init `step`)."""

    def __init__(self, projector, angles: torch.Tensor, iterations: int = 10, step: float = 1e-4,
                 use_checkpoint: bool = True) -> None:
        super().__init__()
        self.projector = projector
        self.register_buffer("angles", torch.as_tensor(angles, dtype=torch.float32), persistent=False)
        self.steps = nn.Parameter(torch.full((iterations,), float(step)))
        self.use_checkpoint = use_checkpoint

    def _iterate(self, x: torch.Tensor, y: torch.Tensor, a: torch.Tensor) -> torch.Tensor:
        r = self.projector.projection(x, self.angles) - y
        return x - a * self.projector.projection_t(r, self.angles)

    def forward(self, y: torch.Tensor, x0: Optional[torch.Tensor] = None) -> torch.Tensor:
        W, H, D = self.projector.image_size
        x = x0 if x0 is not None else torch.zeros((y.shape[0], y.shape[1], D, H, W), device=y.device)
        for k in range(self.steps.numel()):
            if self.use_checkpoint and torch.is_grad_enabled():
                x = checkpoint(self._iterate, x, y, self.steps[k], use_reentrant=False)
            else:
                x = self._iterate(x, y, self.steps[k])
        return x
