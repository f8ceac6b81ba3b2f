"""Loader for the native extension and the plan cache.

The native extension (``xrecon_amd/_C*.so`` over ``_lib/libxrecon_amd.so``) is the only
compute path: if it is missing or cannot be loaded, every operator raises.  GPU tensors
run the gfx950 kernels through device plans; CPU tensors run the library's host executor
(the same per-thread bodies on the host's cores) through host plans, as the reference
runs its CPU kernels for CPU tensors.  There is no Python compute path and no fallback
from one device to the other.
"""
from __future__ import annotations

import collections
import threading
import weakref
from typing import Sequence

import torch

OP_PROJECTION = 0
OP_PROJECTION_T = 1
OP_BACKPROJECTION = 2
OP_BACKPROJECTION_W = 3
OP_BACKPROJECTION_T = 4
OP_BACKPROJECTION_W_T = 5

# Adjoint pairs: the backward of each operator is its transpose
# (xrecon/projector/functions.py:203-302).
ADJOINT = {
    OP_PROJECTION: OP_PROJECTION_T,
    OP_PROJECTION_T: OP_PROJECTION,
    OP_BACKPROJECTION: OP_BACKPROJECTION_T,
    OP_BACKPROJECTION_T: OP_BACKPROJECTION,
    OP_BACKPROJECTION_W: OP_BACKPROJECTION_W_T,
    OP_BACKPROJECTION_W_T: OP_BACKPROJECTION_W,
}

_C = None
_lock = threading.Lock()


def native():
    """The native extension module; raises ImportError when it is not built."""
    global _C
    if _C is None:
        try:
            from . import _C as mod  # noqa: PLC0415
        except ImportError as exc:  # pragma: no cover - exercised only without a build
            raise ImportError(
                "xrecon_amd native extension is not built or cannot be loaded; "
                "run `python -m xrecon_amd.build` (needs hipcc for gfx950)") from exc
        _C = mod
    return _C


def current_stream(t: torch.Tensor) -> int:
    """torch's current HIP stream of t's device (0 for CPU tensors, which need none)."""
    if not t.is_cuda:
        return 0
    return torch.cuda.current_stream(t.device).cuda_stream


class PlanCache:
    """Plans (device geometry tables) keyed by the identity and version counter of the
    options and angles tensors, validated through weak references, with a second level
    keyed by content for host-resident tensors.  A plan is built on the host from the
    angle values, so a miss on device-resident angles costs one device->host copy."""

    def __init__(self, maxsize: int = 64) -> None:
        self.maxsize = maxsize
        self._by_id: "collections.OrderedDict" = collections.OrderedDict()
        self._by_val: "collections.OrderedDict" = collections.OrderedDict()

    def _insert(self, table, key, value):
        table[key] = value
        table.move_to_end(key)
        while len(table) > self.maxsize:
            table.popitem(last=False)

    def get(self, options, angles: torch.Tensor, device: int, is2d: bool):
        okey = tuple(float(v) for v in options) if not isinstance(options, torch.Tensor) else None
        id_key = (id(options) if okey is None else okey, None if okey is not None else options._version,
                  id(angles), angles._version, device, is2d)
        with _lock:
            ent = self._by_id.get(id_key)
            if ent is not None:
                wo, wa, plan = ent
                if (wo is None or wo() is options) and wa() is angles:
                    self._by_id.move_to_end(id_key)
                    return plan
        opts = okey if okey is not None else tuple(float(v) for v in options.detach().cpu().tolist())
        ang_host = angles.detach().to("cpu", torch.float32).contiguous()
        val_key = (opts, ang_host.numpy().tobytes(), device, is2d)
        with _lock:
            plan = self._by_val.get(val_key)
        if plan is None:
            if device < 0:
                plan = native().HostPlan(list(opts), ang_host, is2d)
            else:
                plan = native().Plan(list(opts), ang_host, device, is2d)
            with _lock:
                self._insert(self._by_val, val_key, plan)
        with _lock:
            self._insert(self._by_id, id_key,
                         (None if okey is not None else weakref.ref(options), weakref.ref(angles), plan))
        return plan

    def clear(self) -> None:
        with _lock:
            self._by_id.clear()
            self._by_val.clear()


PLANS = PlanCache()


def get_plan(options, angles: torch.Tensor, ref: torch.Tensor, is2d: bool):
    """The plan for `ref`'s device: a device plan for GPU tensors, a host plan (device
    key -1) for CPU tensors."""
    if angles.dim() != 1:
        raise ValueError(f"Expected angles shape [V], got {tuple(angles.shape)}")
    return PLANS.get(options, angles, ref.get_device() if ref.is_cuda else -1, is2d)


def options_tuple(values: Sequence[float]):
    return tuple(float(torch.tensor(v, dtype=torch.float32).item()) for v in values)
