"""Geometry value objects: the same public fields and conventions as the reference's
``xrecon/projector/geometry.py`` so user code that builds a ``Volume3d`` + ``ConeBeam3d``
(or the 2D pair) works unchanged.

Conventions kept from the reference (they decide what the kernels compute):

* shifts of volumes are given in pixels and converted to mm by the projector modules
  (``projector3d.py:91``);
* the *effective* detector spacing is the virtual (iso-plane) pitch for flat detectors
  and the angular pitch for arc detectors: 3D arc ``du / SDD`` (``geometry.py:199-203``),
  2D arc ``2 atan(du / (2 SDD))`` (``geometry.py:106-109``);
* ``scan_id`` is 0 parallel, 1 flat, 2 arc (``include/types.h:7-11``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple, Union

Number = Union[int, float]
PairLike = Union[Number, Tuple[Number, Number], List[Number]]
TripleLike = Union[Number, Tuple[Number, Number, Number], List[Number]]

_SCAN_IDS = {"parallel": 0, "flat": 1, "arc": 2}


def _expand(value, n: int, name: str):
    if isinstance(value, (tuple, list)):
        if len(value) != n:
            raise ValueError(f"{name} should be a scalar or a {n}-tuple/list")
        return tuple(value)
    return (value,) * n


def to_2tuple(value: PairLike, name: str) -> Tuple[Number, Number]:
    return _expand(value, 2, name)


def to_3tuple(value: TripleLike, name: str) -> Tuple[Number, Number, Number]:
    return _expand(value, 3, name)


def _freeze(obj, **fields):
    for k, v in fields.items():
        object.__setattr__(obj, k, v)


def _detector_type(kind: str) -> str:
    kind = kind.lower().strip()
    if kind not in ("flat", "arc"):
        raise ValueError("detector_type should be 'flat' or 'arc'")
    return kind


@dataclass(frozen=True)
class Volume2d:
    size: Tuple[int, int]
    spacing: Tuple[float, float]
    shift: Tuple[float, float] = (0.0, 0.0)

    def __init__(self, size: PairLike, spacing: PairLike, shift: PairLike = 0.0) -> None:
        _freeze(self, size=tuple(int(v) for v in to_2tuple(size, "size")),
                spacing=tuple(float(v) for v in to_2tuple(spacing, "spacing")),
                shift=tuple(float(v) for v in to_2tuple(shift, "shift")))


@dataclass(frozen=True)
class Detector2d:
    count: int
    spacing: float
    shift: float = 0.0

    def __init__(self, count: int, spacing: float, shift: float = 0.0) -> None:
        _freeze(self, count=int(count), spacing=float(spacing), shift=float(shift))


@dataclass(frozen=True)
class ParallelBeam2d:
    detector: Detector2d
    scan_type: str = "parallel"
    iso_source: float = 0.0
    source_detector: float = 0.0

    @property
    def scan_id(self) -> int:
        return 0

    @property
    def effective_det_spacing(self) -> float:
        return self.detector.spacing

    @property
    def effective_det_shift(self) -> float:
        return self.detector.shift * self.effective_det_spacing


@dataclass(frozen=True)
class FanBeam2d:
    detector: Detector2d
    iso_source: float
    source_detector: float
    detector_type: str = "flat"

    def __init__(self, detector: Detector2d, iso_source: float, source_detector: float,
                 detector_type: str = "flat") -> None:
        _freeze(self, detector=detector, iso_source=float(iso_source),
                source_detector=float(source_detector), detector_type=_detector_type(detector_type))

    @property
    def scan_type(self) -> str:
        return self.detector_type

    @property
    def scan_id(self) -> int:
        return _SCAN_IDS[self.detector_type]

    @property
    def effective_det_spacing(self) -> float:
        if self.detector_type == "flat":
            return self.detector.spacing * self.iso_source / self.source_detector
        return 2.0 * math.atan(self.detector.spacing / (2.0 * self.source_detector))

    @property
    def effective_det_shift(self) -> float:
        return self.detector.shift * self.effective_det_spacing


Geometry2d = Union[ParallelBeam2d, FanBeam2d]


def _scan_name(scan_type: str) -> str:
    scan_type = scan_type.lower().strip()
    return "parallel" if scan_type == "para" else scan_type


def make_geometry2d(scan_type: str, num_det: int, det_size: float, iso_source: float,
                    source_detector: float, binshift: float = 0.0) -> Geometry2d:
    scan_type = _scan_name(scan_type)
    detector = Detector2d(num_det, det_size, binshift)
    if scan_type == "parallel":
        return ParallelBeam2d(detector)
    if scan_type in ("flat", "arc"):
        return FanBeam2d(detector, iso_source, source_detector, scan_type)
    raise ValueError(f"Unsupported scan_type: {scan_type}")


@dataclass(frozen=True)
class Volume3d:
    size: Tuple[int, int, int]
    spacing: Tuple[float, float, float]
    shift: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    def __init__(self, size: TripleLike, spacing: TripleLike, shift: TripleLike = 0.0) -> None:
        _freeze(self, size=tuple(int(v) for v in to_3tuple(size, "size")),
                spacing=tuple(float(v) for v in to_3tuple(spacing, "spacing")),
                shift=tuple(float(v) for v in to_3tuple(shift, "shift")))


@dataclass(frozen=True)
class Detector3d:
    count: Tuple[int, int]
    spacing: Tuple[float, float]
    shift: Tuple[float, float] = (0.0, 0.0)

    def __init__(self, count: PairLike, spacing: PairLike, shift: PairLike = 0.0) -> None:
        _freeze(self, count=tuple(int(v) for v in to_2tuple(count, "count")),
                spacing=tuple(float(v) for v in to_2tuple(spacing, "spacing")),
                shift=tuple(float(v) for v in to_2tuple(shift, "shift")))


@dataclass(frozen=True)
class ConeBeam3d:
    detector: Detector3d
    iso_source: float
    source_detector: float
    detector_type: str = "flat"

    def __init__(self, detector: Detector3d, iso_source: float, source_detector: float,
                 detector_type: str = "flat") -> None:
        _freeze(self, detector=detector, iso_source=float(iso_source),
                source_detector=float(source_detector), detector_type=_detector_type(detector_type))

    @property
    def scan_type(self) -> str:
        return self.detector_type

    @property
    def scan_id(self) -> int:
        return _SCAN_IDS[self.detector_type]

    @property
    def effective_det_spacing(self) -> Tuple[float, float]:
        du, dv = self.detector.spacing
        if self.detector_type == "flat":
            return du * self.iso_source / self.source_detector, dv
        return du / self.source_detector, dv

    @property
    def effective_det_shift(self) -> Tuple[float, float]:
        du, dv = self.effective_det_spacing
        shu, shv = self.detector.shift
        return shu * du, shv * dv


@dataclass(frozen=True)
class ParallelBeam3d:
    detector: Detector3d
    scan_type: str = "parallel"

    @property
    def scan_id(self) -> int:
        return 0

    @property
    def effective_det_spacing(self) -> Tuple[float, float]:
        return self.detector.spacing

    @property
    def effective_det_shift(self) -> Tuple[float, float]:
        du, dv = self.detector.spacing
        shu, shv = self.detector.shift
        return shu * du, shv * dv


Geometry3d = Union[ParallelBeam3d, ConeBeam3d]


def make_geometry3d(scan_type: str, num_det: PairLike, det_size: PairLike, iso_source: float,
                    source_detector: float, binshift: PairLike = 0.0) -> Geometry3d:
    scan_type = _scan_name(scan_type)
    detector = Detector3d(num_det, det_size, binshift)
    if scan_type == "parallel":
        return ParallelBeam3d(detector)
    if scan_type in ("flat", "arc"):
        return ConeBeam3d(detector, iso_source, source_detector, scan_type)
    raise ValueError(f"Unsupported 3D scan_type: {scan_type}")
