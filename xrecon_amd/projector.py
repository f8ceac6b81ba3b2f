"""``Projector2d``: drop-in for the reference's ``xrecon.Projector2d``
(``xrecon/projector/projector.py:21-313``): [B, C, H, W] images, [B, C, V, Nu] sinograms,
the 12-entry options buffer, cosine weight, filters, Parker weights and FBP.  2D
operators run on the same gfx950 kernels as 3D, on a one-slice volume seen by a one-row
detector with the 2D conventions (effective pitch, no depth magnification).
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple, Union

import torch
import torch.nn as nn
from torch import Tensor

from . import _backend as be
from .filters import correlate_rows, filter_generate
from .filters import fused_fdk_ok as _fused_fdk_ok
from .functions import LinearOp
from .geometry import FanBeam2d, Geometry2d, ParallelBeam2d, Volume2d, make_geometry2d

_SCAN_ALIASES = {"para": "parallel", "parallel": "parallel", "flat": "flat", "arc": "arc"}


class Projector2d(nn.Module):
    """Distance-driven 2D CT projector (parallel, fan-flat, fan-arc).

    Shapes:
        image:  [B, C, H, W]
        angles: [V]
        sino:   [B, C, V, Nu]
    """

    def __init__(
        self,
        image_size: Union[int, Tuple[int, int], List[int], Volume2d],
        num_det: Optional[int] = None,
        pix_size: Union[float, Tuple[float, float], List[float], Geometry2d, None] = None,
        det_size: Optional[float] = None,
        iso_source: float = 0.0,
        source_detector: float = 0.0,
        pixshift: Union[float, Tuple[float, float], List[float]] = 0.0,
        binshift: float = 0.0,
        scan_type: str = "flat",
        filter_type: str = "ramp",
        trainable: bool = False,
        dtype=None,
    ) -> None:
        super().__init__()
        if isinstance(image_size, Volume2d):
            volume = image_size
            if isinstance(num_det, (FanBeam2d, ParallelBeam2d)) and pix_size is None:
                pix_size = num_det
            if not isinstance(pix_size, (FanBeam2d, ParallelBeam2d)):
                raise TypeError("When image_size is Volume2d, pix_size should be a 2D geometry object")
            geometry = pix_size
        else:
            if num_det is None or pix_size is None or det_size is None:
                raise TypeError("num_det, pix_size and det_size are required with the legacy constructor")
            volume = Volume2d(image_size, pix_size, pixshift)
            geometry = make_geometry2d(scan_type, num_det, det_size, iso_source, source_detector, binshift)

        raw = geometry.scan_type.lower().strip()
        scan_type = _SCAN_ALIASES.get(raw, raw)
        if scan_type not in ("parallel", "flat", "arc"):
            raise ValueError(f"Unsupported scan_type: {scan_type}")

        W, H = volume.size
        sx, sy = volume.spacing
        shx, shy = volume.shift
        self.image_size = (W, H)
        self.pix_size = (sx, sy)
        self.volume = volume
        self.geometry = geometry
        self.num_det = geometry.detector.count
        self.scan_type = scan_type
        self.filter_type = filter_type.lower().strip()
        self.iso_source = float(geometry.iso_source)
        self.source_detector = float(geometry.source_detector)
        self._default_dtype = dtype or torch.float32
        self.det_size = geometry.effective_det_spacing
        self.binshift = geometry.effective_det_shift
        self.pixshift = (shx * sx, shy * sy)

        values = [W, H, self.num_det, sx, sy, self.det_size, self.iso_source, self.source_detector,
                  self.pixshift[0], self.pixshift[1], self.binshift, geometry.scan_id]
        options = torch.tensor(values, dtype=self._default_dtype)
        self.register_buffer("options", options, persistent=False)
        self._host_options = tuple(float(v) for v in options.float().tolist())

        weight = self._make_cosine_weight(dtype=self._default_dtype)
        if trainable:
            self.weight = nn.parameter.Parameter(weight.clone())
        else:
            self.register_buffer("weight", weight, persistent=False)
        filt = self._filter_generate(filter_type=self.filter_type, dtype=self._default_dtype)
        if trainable:
            self.filter = nn.parameter.Parameter(filt.clone())
        else:
            self.register_buffer("filter", filt, persistent=False)

    def plan(self, angles: Tensor, device=None):
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        return be.get_plan(self._host_options, torch.as_tensor(angles), torch.empty(0, device=device), True)

    def _run(self, op: int, x: Tensor, angles: Tensor) -> Tensor:
        plan = be.get_plan(self._host_options, torch.as_tensor(angles), x, True)
        return LinearOp.apply(x, plan, op)

    # ---------- public API ----------
    def forward(self, image: Tensor, angles: Tensor) -> Tensor:
        return self.projection(image, angles)

    def projection(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._run(be.OP_PROJECTION, image, angles)

    def backprojection(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino_angles(sino, angles)
        return self._run(be.OP_BACKPROJECTION, sino, angles)

    def backprojection_weighted(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino_angles(sino, angles)
        return self._run(be.OP_BACKPROJECTION_W, sino, angles)

    def projection_t(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino_angles(sino, angles)
        return self._run(be.OP_PROJECTION_T, sino, angles)

    def backprojection_t(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._run(be.OP_BACKPROJECTION_T, image, angles)

    def backprojection_weighted_t(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._run(be.OP_BACKPROJECTION_W_T, image, angles)

    projection_transpose = projection_t
    backprojection_transpose = backprojection_t
    weighted_backprojection_transpose = backprojection_weighted_t

    def system_matrix(self, angles: Tensor, device=None) -> Tensor:
        """Sparse COO matrix of the projection; built on the host (as the reference), or
        emitted on ``device`` (a GPU) with the same triplets."""
        from .sysmat import system_matrix2d  # noqa: PLC0415

        angles = torch.as_tensor(angles, dtype=torch.float32, device="cpu").contiguous()
        return system_matrix2d(self._host_options, angles, device)

    get_sys_matrix = system_matrix

    def filtered_backprojection(self, sino: Tensor, angles: Tensor, filter: Optional[Tensor] = None,
                                filter_type: Optional[str] = None, redundant: Optional[bool] = None,
                                parker: bool = False) -> Tensor:
        """FBP (projector.py:176-206)."""
        self._check_sino_angles(sino, angles)
        if filter is not None:
            filt = filter
        elif filter_type is None or filter_type.lower() == self.filter_type:
            filt = self.filter
        else:
            filt = self._filter_generate(filter_type=filter_type.lower(), dtype=sino.dtype).to(sino.device)
        if parker:
            sino = self.apply_parker_weight(sino, angles)
            if redundant is None:
                redundant = False
        elif redundant is None:
            redundant = True
        scale = self._mean_angle_step(angles).to(device=sino.device, dtype=sino.dtype) * (0.5 if redundant else 1.0)
        if _fused_fdk_ok(sino, filt, self.weight):
            # one native pass: cosine weight, row filter (FP64 sums), weighted backprojection
            plan = be.get_plan(self._host_options, torch.as_tensor(angles), sino, True)
            taps = (filt.reshape(-1).to(device=sino.device, dtype=torch.float32) * scale).contiguous()
            w = self.weight.reshape(-1).to(device=sino.device, dtype=torch.float32).contiguous()
            return plan.fdk(sino.contiguous(), taps, w, be.current_stream(sino))
        sino = sino * self.weight.to(device=sino.device, dtype=sino.dtype)
        filtered = correlate_rows(sino, filt.to(device=sino.device, dtype=sino.dtype))
        recon = self.backprojection_weighted(filtered, angles)
        return recon * scale

    def apply_parker_weight(self, sino: Tensor, angles: Tensor) -> Tensor:
        return sino * self.parker_weight(angles).to(device=sino.device, dtype=sino.dtype)

    def parker_weight(self, angles: Tensor) -> Tensor:
        """Parker short-scan weights (projector.py:211-249), [1, 1, V, Nu]."""
        angles = torch.as_tensor(angles)
        V, Nu = angles.numel(), self.num_det
        ones = torch.ones((1, 1, V, Nu), device=angles.device, dtype=angles.dtype)
        if self.scan_type not in ("flat", "arc") or V < 2:
            return ones
        beta = angles - angles[0]
        scan_range = beta[-1].abs()
        gamma = self._fan_angles(device=angles.device, dtype=angles.dtype)
        gamma_max = gamma.abs().max()
        if scan_range >= (2 * math.pi - 1e-4):
            return ones
        if scan_range < math.pi + 2.0 * gamma_max - self._mean_angle_step(angles):
            raise ValueError("Parker weighting requires a short-scan range of at least "
                             f"pi + 2 * gamma_max, got {scan_range.item():.6f} rad")
        overscan = torch.clamp((scan_range - math.pi) / 2.0, min=gamma_max)
        eps = torch.finfo(angles.dtype).eps
        left = 2.0 * (overscan - gamma)
        right = scan_range - 2.0 * (overscan + gamma)
        b2 = beta[:, None]
        w = torch.ones((V, Nu), device=angles.device, dtype=angles.dtype)
        w_left = torch.sin(0.5 * math.pi * b2 / torch.clamp(left, min=eps)[None, :]).pow(2)
        w_right = torch.sin(0.5 * math.pi * (scan_range - b2) / torch.clamp(scan_range - right, min=eps)[None, :]).pow(2)
        w = torch.where(b2 < left[None, :], w_left, w)
        w = torch.where(b2 > right[None, :], w_right, w)
        return w.clamp_(0.0, 1.0).view(1, 1, V, Nu)

    def _filter_generate(self, filter_type: str, dtype=None) -> Tensor:
        return filter_generate(self.num_det, self.det_size, filter_type, dtype)[None, None, None, :]

    def _detector_coords(self, device=None, dtype=None) -> Tensor:
        u = torch.arange(-self.num_det / 2 + 0.5, self.num_det / 2, 1.0, device=device, dtype=dtype or torch.float32)
        return u * self.det_size + self.binshift

    def _make_cosine_weight(self, dtype=None) -> Tensor:
        """projector.py:258-283."""
        s = self._detector_coords(dtype=dtype or torch.float32)
        if self.scan_type == "flat":
            w = self.iso_source / torch.sqrt(s * s + self.iso_source * self.iso_source)
        elif self.scan_type == "arc":
            w = torch.cos(s)
        else:
            w = torch.ones_like(s)
        return w.view(1, 1, 1, self.num_det)

    def _fan_angles(self, device=None, dtype=None) -> Tensor:
        s = self._detector_coords(device=device, dtype=dtype)
        if self.scan_type == "arc":
            return s
        if self.scan_type == "flat":
            return torch.atan2(s, torch.as_tensor(self.iso_source, device=device, dtype=s.dtype))
        return torch.zeros_like(s)

    def _check_image(self, image: Tensor) -> None:
        if image.ndim != 4:
            raise ValueError(f"Expected image shape [B, C, H, W], got {tuple(image.shape)}")
        W, H = self.image_size
        if tuple(image.shape[2:]) != (H, W):
            raise ValueError(f"Expected input image size {(H, W)}, but got {tuple(image.shape[2:])}")

    def _check_sino_angles(self, sino: Tensor, angles: Tensor) -> None:
        if sino.ndim != 4:
            raise ValueError(f"Expected sino shape [B, C, V, Nu], got {tuple(sino.shape)}")
        if angles.ndim != 1:
            raise ValueError(f"Expected angles shape [V], got {tuple(angles.shape)}")
        if sino.shape[2] != angles.numel() or sino.shape[3] != self.num_det:
            raise ValueError(f"Expected sino shape [B, C, {angles.numel()}, {self.num_det}], got {tuple(sino.shape)}")

    @staticmethod
    def _prepare_angles(angles: Tensor, ref: Tensor) -> Tensor:
        return angles.to(device=ref.device, dtype=ref.dtype).contiguous()

    @staticmethod
    def _mean_angle_step(angles: Tensor) -> Tensor:
        angles = torch.as_tensor(angles)
        if angles.numel() < 2:
            return angles.new_tensor(0.0)
        wrap = (torch.diff(angles) + math.pi) % (2 * math.pi) - math.pi
        return wrap.abs().mean()


projector2d = Projector2d
