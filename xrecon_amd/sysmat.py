"""System matrices (sparse COO of the forward projector), as the reference's
``xrecon.projector.functions.system_matrix2d/3d`` return them
(``src/api/torch_binding.cpp:353-405``): a coalesced ``torch.sparse_coo_tensor`` of
shape [V*Nv*Nu, D*H*W] (3D) or [V*Nu, H*W] (2D) with the same index and sparsity pattern
as the reference's CPU builder.

Without ``device`` the matrix is built on the host (the reference's behaviour, CPU
tensor out).  With ``device`` (a GPU) the same triplets are emitted there, bit for bit,
and the result is a GPU sparse tensor; the reference has no GPU builder.
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from . import _backend as be


def _opts(options) -> torch.Tensor:
    if isinstance(options, torch.Tensor):
        return options.detach().to("cpu", torch.float32).contiguous()
    return torch.tensor([float(v) for v in options], dtype=torch.float32)


def _angles(angles) -> torch.Tensor:
    return torch.as_tensor(angles, dtype=torch.float32).detach().cpu().contiguous()


def _shape(options, nview: int, is2d: bool):
    o = [float(v) for v in _opts(options).tolist()]
    if is2d:
        return (nview * int(o[2]), int(o[0]) * int(o[1]))
    return (nview * int(o[4]) * int(o[3]), int(o[2]) * int(o[1]) * int(o[0]))


def triplets(options: Sequence[float], angles, is2d: bool = False, device=None):
    """(rows, cols, values) before coalescing, in the traversal order of the reference's
    builder: from the host builder, or emitted on `device`."""
    dev = -1 if device is None else torch.device(device).index or 0
    return tuple(be.native().system_matrix_triplets(_opts(options), _angles(angles), is2d, dev))


def _build(options, angles, is2d: bool, device: Optional[torch.device]):
    if device is None:
        fn = be.native().system_matrix2d if is2d else be.native().system_matrix3d
        return fn(_opts(options), _angles(angles))
    rows, cols, vals = triplets(options, angles, is2d, device)
    shape = _shape(options, _angles(angles).numel(), is2d)
    return torch.sparse_coo_tensor(torch.stack([rows, cols]), vals, shape).coalesce()


def system_matrix3d(options: Sequence[float], angles, device=None) -> torch.Tensor:
    return _build(options, angles, False, device)


def system_matrix2d(options: Sequence[float], angles, device=None) -> torch.Tensor:
    return _build(options, angles, True, device)
