"""``Projector3d``: drop-in for the reference's ``xrecon.Projector3d``
(``xrecon/projector/projector3d.py:29-287``): same constructor, buffers (``options``,
``weight``, ``filter``), shapes ([B, C, D, H, W] volumes, [B, C, V, Nv, Nu] sinograms,
[V] angles) and methods.  The operators run on hand-written gfx950 kernels through the
native extension; plans (per-view FP64 geometry tables) are cached per angle tensor so a
training or reconstruction loop pays for them once.
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
from .geometry import ConeBeam3d, Geometry3d, ParallelBeam3d, Volume3d, make_geometry3d


class Projector3d(nn.Module):
    """Distance-driven circular cone-beam (flat / arc) or parallel-beam 3D projector.

    Shapes:
        image:  [B, C, D, H, W]
        angles: [V]
        sino:   [B, C, V, Nv, Nu]
    """

    def __init__(
        self,
        image_size: Union[int, Tuple[int, int, int], List[int], Volume3d],
        num_det: Optional[Union[int, Tuple[int, int], List[int]]] = None,
        pix_size: Union[float, Tuple[float, float, float], List[float], Geometry3d, None] = None,
        det_size: Optional[Union[float, Tuple[float, float], List[float]]] = None,
        iso_source: float = 0.0,
        source_detector: float = 0.0,
        pixshift: Union[float, Tuple[float, float, float], List[float]] = 0.0,
        binshift: Union[float, Tuple[float, float], List[float]] = 0.0,
        scan_type: str = "flat",
        filter_type: str = "ramp",
        trainable: bool = False,
        dtype=None,
    ) -> None:
        super().__init__()
        if isinstance(image_size, Volume3d):
            volume = image_size
            if isinstance(num_det, (ConeBeam3d, ParallelBeam3d)) and pix_size is None:
                pix_size = num_det
            if not isinstance(pix_size, (ConeBeam3d, ParallelBeam3d)):
                raise TypeError("When image_size is Volume3d, pix_size should be a Geometry3d object")
            geometry = pix_size
        else:
            if num_det is None or pix_size is None or det_size is None:
                raise TypeError("num_det, pix_size and det_size are required with the legacy constructor")
            volume = Volume3d(image_size, pix_size, pixshift)
            geometry = make_geometry3d(scan_type, num_det, det_size, iso_source, source_detector, binshift)

        scan_type = geometry.scan_type.lower().strip()
        if scan_type not in ("parallel", "flat", "arc"):
            raise ValueError(f"Unsupported 3D scan_type: {scan_type}")

        W, H, D = volume.size
        dx, dy, dz = volume.spacing
        shx, shy, shz = volume.shift
        Nu, Nv = geometry.detector.count
        du, dv = geometry.effective_det_spacing
        shu, shv = geometry.effective_det_shift
        raw_du, raw_dv = geometry.detector.spacing
        raw_shu, raw_shv = geometry.detector.shift
        # The kernels take the flat detector at its physical pitch and scale it to the iso
        # plane themselves (projector3d.py:81-82); arc and parallel take the effective one.
        kernel_du = raw_du if scan_type == "flat" else du
        kernel_shu = raw_shu * raw_du if scan_type == "flat" else shu

        self.volume = volume
        self.geometry = geometry
        self.image_size = (W, H, D)
        self.pix_size = (dx, dy, dz)
        self.num_det = (Nu, Nv)
        self.det_size = (du, dv)
        self.binshift = (shu, shv)
        self.pixshift = (shx * dx, shy * dy, shz * dz)
        self.scan_type = scan_type
        self.filter_type = filter_type.lower().strip()
        self.iso_source = float(getattr(geometry, "iso_source", 0.0))
        self.source_detector = float(getattr(geometry, "source_detector", 0.0))
        self._default_dtype = dtype or torch.float32

        values = [W, H, D, Nu, Nv, dx, dy, dz, kernel_du, raw_dv, self.iso_source, self.source_detector,
                  self.pixshift[0], self.pixshift[1], self.pixshift[2], kernel_shu, raw_shv * raw_dv,
                  geometry.scan_id]
        options = torch.tensor(values, dtype=self._default_dtype)
        self.register_buffer("options", options, persistent=False)
        # Host copy of the options (float32-rounded, as the kernels receive them): plans are
        # built from it without reading the device buffer.
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

    # ------------------------------------------------------------------ plans
    def plan(self, angles: Tensor, device=None):
        """The native plan for `angles` on `device` (default: current CUDA device)."""
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        ref = torch.empty(0, device=device)
        return be.get_plan(self._host_options, torch.as_tensor(angles), ref, False)

    def _run(self, op: int, x: Tensor, angles: Tensor) -> Tensor:
        angles = torch.as_tensor(angles)
        plan = be.get_plan(self._host_options, angles, x, False)
        return LinearOp.apply(x, plan, op)

    # ------------------------------------------------------------------ operators
    def forward(self, image: Tensor, angles: Tensor) -> Tensor:
        return self.projection(image, angles)

    def projection(self, image: Tensor, angles: Tensor) -> Tensor:
        self._check_image(image)
        return self._run(be.OP_PROJECTION, image, angles)

    def projection_t(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino_angles(sino, angles)
        return self._run(be.OP_PROJECTION_T, sino, angles)

    def backprojection(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino_angles(sino, angles)
        return self._run(be.OP_BACKPROJECTION, sino, angles)

    def backprojection_weighted(self, sino: Tensor, angles: Tensor) -> Tensor:
        self._check_sino_angles(sino, angles)
        return self._run(be.OP_BACKPROJECTION_W, sino, angles)

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
        from .sysmat import system_matrix3d  # noqa: PLC0415

        angles = torch.as_tensor(angles, dtype=torch.float32, device="cpu").contiguous()
        return system_matrix3d(self._host_options, angles, device)

    get_sys_matrix = system_matrix

    def filtered_backprojection(self, sino: Tensor, angles: Tensor, filter: Optional[Tensor] = None,
                                filter_type: Optional[str] = None, redundant: Optional[bool] = None,
                                parker: bool = False) -> Tensor:
        """FDK (projector3d.py:165-195): weights, row ramp filter, FDK-weighted
        backprojection, times the mean angular step, halved for a full (redundant) scan."""
        self._check_sino_angles(sino, angles)
        if filter is not None:
            filt = filter
        elif filter_type is None or filter_type.lower() == self.filter_type:
            filt = self.filter
        else:
            filt = self._filter_generate(filter_type.lower(), dtype=sino.dtype).to(sino.device)
        if parker:
            sino = sino * self.parker_weight(angles).to(device=sino.device, dtype=sino.dtype)
            if redundant is None:
                redundant = False
        elif redundant is None:
            redundant = True
        scale = self._mean_angle_step(angles).to(device=sino.device, dtype=sino.dtype) * (0.5 if redundant else 1.0)
        if _fused_fdk_ok(sino, filt, self.weight):
            # one native pass: cosine weight, row filter (FP64 sums), weighted backprojection
            plan = be.get_plan(self._host_options, torch.as_tensor(angles), sino, False)
            taps = (filt.reshape(-1).to(device=sino.device, dtype=torch.float32) * scale).contiguous()
            w = self.weight.reshape(-1).to(device=sino.device, dtype=torch.float32).contiguous()
            return plan.fdk(sino.contiguous(), taps, w, be.current_stream(sino))
        sino = sino * self.weight.to(device=sino.device, dtype=sino.dtype)
        filtered = correlate_rows(sino, filt.to(device=sino.device, dtype=sino.dtype))
        recon = self.backprojection_weighted(filtered, angles)
        return recon * scale

    def parker_weight(self, angles: Tensor) -> Tensor:
        """Parker short-scan weights (projector3d.py:197-227), [1, 1, V, 1, Nu]."""
        angles = torch.as_tensor(angles)
        V, Nu = angles.numel(), self.num_det[0]
        ones = torch.ones((1, 1, V, 1, Nu), device=angles.device, dtype=angles.dtype)
        if V < 2 or self.scan_type == "parallel":
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
        return w.clamp_(0.0, 1.0).view(1, 1, V, 1, Nu)

    # ------------------------------------------------------------------ helpers
    def _filter_generate(self, filter_type: str, dtype=None) -> Tensor:
        return filter_generate(self.num_det[0], self.det_size[0], filter_type, dtype)[None, None, None, None, :]

    def _detector_coords(self, device=None, dtype=None) -> Tensor:
        Nu = self.num_det[0]
        u = torch.arange(-Nu / 2 + 0.5, Nu / 2, 1.0, device=device, dtype=dtype or torch.float32)
        return u * self.det_size[0] + self.binshift[0]

    def _make_cosine_weight(self, dtype=None) -> Tensor:
        """Cosine pre-weight along u only (projector3d.py:233-246)."""
        dtype = dtype or torch.float32
        Nu = self.num_det[0]
        s = self._detector_coords(dtype=dtype)
        if self.scan_type == "parallel":
            return torch.ones((1, 1, 1, 1, Nu), dtype=dtype)
        if self.scan_type == "flat":
            w = self.iso_source / torch.sqrt(s * s + self.iso_source * self.iso_source)
        else:
            w = torch.cos(s)
        return w.view(1, 1, 1, 1, Nu)

    def _fan_angles(self, device=None, dtype=None) -> Tensor:
        s = self._detector_coords(device=device, dtype=dtype)
        if self.scan_type in ("parallel", "arc"):
            return s
        return torch.atan2(s, torch.as_tensor(self.iso_source, device=device, dtype=s.dtype))

    def _check_image(self, image: Tensor) -> None:
        if image.ndim != 5:
            raise ValueError(f"Expected image shape [B, C, D, H, W], got {tuple(image.shape)}")
        W, H, D = self.image_size
        if tuple(image.shape[2:]) != (D, H, W):
            raise ValueError(f"Expected image shape [B, C, {D}, {H}, {W}], got {tuple(image.shape)}")

    def _check_sino_angles(self, sino: Tensor, angles: Tensor) -> None:
        if sino.ndim != 5:
            raise ValueError(f"Expected sino shape [B, C, V, Nv, Nu], got {tuple(sino.shape)}")
        if angles.ndim != 1:
            raise ValueError(f"Expected angles shape [V], got {tuple(angles.shape)}")
        Nu, Nv = self.num_det
        if sino.shape[2] != angles.numel() or sino.shape[3] != Nv or sino.shape[4] != Nu:
            raise ValueError(f"Expected sino shape [B, C, {angles.numel()}, {Nv}, {Nu}], got {tuple(sino.shape)}")

    @staticmethod
    def _prepare_angles(angles: Tensor, ref: Tensor) -> Tensor:
        return angles.to(device=ref.device, dtype=ref.dtype).contiguous()

    @staticmethod
    def _mean_angle_step(angles: Tensor) -> Tensor:
        """Mean wrapped angular step (projector3d.py:282-287)."""
        angles = torch.as_tensor(angles)
        if angles.numel() < 2:
            return angles.new_tensor(0.0)
        wrap = (torch.diff(angles) + math.pi) % (2 * math.pi) - math.pi
        return wrap.abs().mean()


projector3d = Projector3d
