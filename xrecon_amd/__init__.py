"""xrecon_amd — MI355X-native (gfx950) distance-driven CT projector core.

A drop-in for the reference's ``xrecon.Projector2d`` / ``xrecon.Projector3d`` path:
projection, backprojection, FBP/FDK-weighted backprojection, their transposes and
filtered backprojection, for parallel, flat and arc geometries, on [B, C, D, H, W]
volumes and [B, C, V, Nv, Nu] sinograms.  Compute runs in hand-written HIP kernels
(``xrecon_amd/csrc``) behind the C-ABI in ``include/xrecon_amd.h``.
"""
from .geometry import (ConeBeam3d, Detector2d, Detector3d, FanBeam2d, ParallelBeam2d, ParallelBeam3d, Volume2d,
                       Volume3d, make_geometry2d, make_geometry3d)
from .projector import Projector2d, projector2d
from .projector3d import Projector3d, projector3d

__all__ = [
    "Detector2d",
    "Detector3d",
    "FanBeam2d",
    "ConeBeam3d",
    "ParallelBeam2d",
    "ParallelBeam3d",
    "Projector2d",
    "Projector3d",
    "Volume2d",
    "Volume3d",
    "make_geometry2d",
    "make_geometry3d",
    "projector2d",
    "projector3d",
]

__version__ = "0.1.0"
