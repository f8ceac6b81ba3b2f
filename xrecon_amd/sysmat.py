"""System matrices (sparse COO of the forward projector), as the reference's
``xrecon.projector.functions.system_matrix2d/3d`` return them
(``src/api/torch_binding.cpp:353-405``): a coalesced ``torch.sparse_coo_tensor`` of
shape [V*Nv*Nu, D*H*W] (3D) or [V*Nu, H*W] (2D), built on the host with the same index
and sparsity pattern as the reference (see ``csrc/sysmat.cpp``)."""
from __future__ import annotations

from typing import Sequence

import torch

from . import _backend as be


def _opts(options) -> torch.Tensor:
    if isinstance(options, torch.Tensor):
        return options.detach().to("cpu", torch.float32).contiguous()
    return torch.tensor([float(v) for v in options], dtype=torch.float32)


def system_matrix3d(options: Sequence[float], angles) -> torch.Tensor:
    ang = torch.as_tensor(angles, dtype=torch.float32).detach().cpu().contiguous()
    return be.native().system_matrix3d(_opts(options), ang)


def system_matrix2d(options: Sequence[float], angles) -> torch.Tensor:
    ang = torch.as_tensor(angles, dtype=torch.float32).detach().cpu().contiguous()
    return be.native().system_matrix2d(_opts(options), ang)
