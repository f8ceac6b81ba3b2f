"""Autotuner for the gather kernels' tile shapes.

Two knobs exist, one per kernel family (``Plan.kernel_kind(op)``):

* cell gathers (A, B^T, B_W^T): ``tile_rows`` in {1, 2, 4, 8, 16} -- detector rows per
  thread.  More rows amortise each line's column overlaps and share row boundaries,
  but cost registers and therefore resident waves.  Tiles of 4 rows and more are also
  tried with CELL_LDS and CELL_DIRECT: each workgroup's table lookups served from
  blocks staged in LDS, or read from device memory (the same results either way).
* voxel gathers (B, B_W, A^T): ``voxel_chunk`` in {1, 4, 8, 16} -- slices per thread --
  and 16 | VOXEL_HALVES, 32 | VOXEL_HALVES: 16 or 32 slices run as two halves that a wave
  skips when none of its lanes reaches the detector in them.  Longer chunks amortise each view's column search,
  shorter ones give more waves; the halves pay on tall volumes.

``tune`` times the gather pass alone (the prefix pass is shared) for every candidate
with HIP events on the caller's stream, takes the median of a few launches, and keeps
the winner on the plan object itself, per (operator, batch): a plan is one geometry and
one angle set, and the choice lives and dies with it.  Tuning launches kernels and
synchronises, so run it before graph capture or timed regions.
"""
from __future__ import annotations

import statistics
from typing import Dict, Tuple

import torch

CELL = 1
VOXEL = 2
TILE_ROWS = (1, 2, 4, 8, 16)
CELL_DIRECT = 0x100  # XR_CELL_DIRECT
CELL_LDS = 0x200  # XR_CELL_LDS
VOXEL_HALVES = 0x100  # XR_VOXEL_HALVES
VOXEL_CHUNKS = (1, 4, 8, 16, 16 | VOXEL_HALVES, 32 | VOXEL_HALVES)


def _table(plan) -> Dict[tuple, Tuple[int, int]]:
    table = getattr(plan, "_xr_tuned", None)
    if table is None:
        table = {}
        plan._xr_tuned = table
    return table


def cached(plan, op: int, nbatch: int):
    return _table(plan).get((op, nbatch))


def candidates(plan, op: int):
    if plan.kernel_kind(op) == CELL:
        return [(r | m, 0) for r in TILE_ROWS for m in ((CELL_DIRECT, CELL_LDS) if r >= 4 else (0,))]
    return [(0, k) for k in VOXEL_CHUNKS]


def tune(plan, op: int, x: torch.Tensor, out: torch.Tensor, ws: torch.Tensor, reps: int = 3,
         verbose: bool = False) -> Tuple[int, int]:
    """Pick (tile_rows, voxel_chunk) for `op` on input `x` (shape fixes the batch).
    Returns the cached choice if this (plan, op, batch) was tuned before."""
    nb = x.shape[0] * x.shape[1]
    key = (op, nb)
    table = _table(plan)
    if key in table:
        return table[key]
    stream = torch.cuda.current_stream(x.device)
    sid = stream.cuda_stream
    plan.execute_into(op, x, out, ws, sid, 0, 0, 1)  # prefix pass, shared by all candidates
    best, best_t = None, float("inf")
    timings = {}
    for rows, chunk in candidates(plan, op):
        plan.execute_into(op, x, out, ws, sid, rows, chunk, 2)  # warm-up
        ts = []
        for _ in range(reps):
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            plan.execute_into(op, x, out, ws, sid, rows, chunk, 2)
            e.record(stream)
            e.synchronize()
            ts.append(s.elapsed_time(e))
        t = statistics.median(ts)
        timings[(rows, chunk)] = t
        if t < best_t:
            best, best_t = (rows, chunk), t
    if verbose:
        print("autotune op", op, {k: round(v, 3) for k, v in timings.items()}, "->", best, flush=True)
    table[key] = best
    return best


def clear(plan) -> None:
    """Forget the choices tuned on `plan`."""
    _table(plan).clear()
