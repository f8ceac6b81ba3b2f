"""Many uneven projector instances in single launches, optionally captured as a graph.

``BatchedProjector3d`` groups up to 32 ``Projector3d`` instances of one geometry family
(different volumes, detectors, view counts and angle sets) and runs one operator for all
of them with one launch per pass (``include/xrecon_amd.h``, ``xr_batch3d_*``).  The
reference has no equivalent: it runs one instance per call.

* ``run(op, inputs)`` — one batched execution (allocates outputs unless given).
* ``tune(op, inputs)`` — picks the launch's tile shape (autotuner, see ``autotune``).
* ``capture(ops, inputs)`` — records a sequence of batched operators into a HIP graph
  (``torch.cuda.CUDAGraph``) over static input/output buffers; ``replay()`` reruns it
  with no host work per launch.
"""
from __future__ import annotations

import statistics
from typing import List, Optional, Sequence

import torch

from . import _backend as be
from .autotune import CELL, TILE_ROWS, VOXEL_CHUNKS


class BatchedProjector3d:
    def __init__(self, projectors: Sequence, angles: Sequence[torch.Tensor], batch_counts: Optional[Sequence[int]] = None,
                 device=None) -> None:
        if len(projectors) != len(angles):
            raise ValueError("one angle set per projector")
        C = be.native()
        if not 1 <= len(projectors) <= C.BATCH_MAX:
            raise ValueError(f"a batch holds 1..{C.BATCH_MAX} projectors")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.projectors = list(projectors)
        self.angles = [torch.as_tensor(a, dtype=torch.float32) for a in angles]
        self.batch_counts = list(batch_counts) if batch_counts is not None else [1] * len(projectors)
        self.plans = [p.plan(a, self.device) for p, a in zip(self.projectors, self.angles)]
        self.batch = C.Batch(self.plans, self.batch_counts)
        self._ws = {}
        self._tiles = {}
        self._graph = None

    # ----------------------------------------------------------------- shapes
    def volume_shape(self, i: int):
        W, H, D = self.projectors[i].image_size
        return (self.batch_counts[i], 1, D, H, W)

    def sino_shape(self, i: int):
        Nu, Nv = self.projectors[i].num_det
        return (self.batch_counts[i], 1, self.angles[i].numel(), Nv, Nu)

    def output_shape(self, op: int, i: int):
        return self.sino_shape(i) if self.plans[0].kernel_kind(op) == CELL else self.volume_shape(i)

    def voxel_views(self) -> int:
        """Sum over instances of Nx*Ny*Nz*V*batch (the unit of work of one operator)."""
        total = 0
        for i, p in enumerate(self.projectors):
            W, H, D = p.image_size
            total += W * H * D * self.angles[i].numel() * self.batch_counts[i]
        return total

    # ----------------------------------------------------------------- execution
    def workspace(self, op: int) -> torch.Tensor:
        n = self.batch.workspace_bytes(op)
        ws = self._ws.get(n)
        if ws is None:
            ws = torch.empty(max(n, 1), dtype=torch.uint8, device=self.device)
            self._ws[n] = ws
        return ws

    def empty_outputs(self, op: int) -> List[torch.Tensor]:
        return [torch.empty(self.output_shape(op, i), device=self.device) for i in range(len(self.plans))]

    def run(self, op: int, inputs: Sequence[torch.Tensor], outputs: Optional[List[torch.Tensor]] = None,
            tiles=None) -> List[torch.Tensor]:
        inputs = [x.contiguous() for x in inputs]
        if outputs is None:
            outputs = self.empty_outputs(op)
        rows, chunk = tiles if tiles is not None else self._tiles.get(op, (0, 0))
        self.batch.execute_into(op, inputs, outputs, self.workspace(op), be.current_stream(inputs[0]), rows, chunk)
        return outputs

    def tune(self, op: int, inputs: Sequence[torch.Tensor], reps: int = 3, verbose: bool = False):
        """Choose the tile shape of the batched launch by timing every candidate."""
        if op in self._tiles:
            return self._tiles[op]
        outs = self.empty_outputs(op)
        cands = [(r, 0) for r in TILE_ROWS] if self.plans[0].kernel_kind(op) == CELL else [(0, k) for k in VOXEL_CHUNKS]
        stream = torch.cuda.current_stream(self.device)
        best, best_t, times = None, float("inf"), {}
        for t in cands:
            self.run(op, inputs, outs, tiles=t)
            ts = []
            for _ in range(reps):
                s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s.record(stream)
                self.run(op, inputs, outs, tiles=t)
                e.record(stream)
                e.synchronize()
                ts.append(s.elapsed_time(e))
            times[t] = statistics.median(ts)
            if times[t] < best_t:
                best, best_t = t, times[t]
        if verbose:
            print("batched autotune op", op, {k: round(v, 3) for k, v in times.items()}, "->", best, flush=True)
        self._tiles[op] = best
        return best

    # ----------------------------------------------------------------- graphs
    def capture(self, ops: Sequence[int], inputs: Sequence[torch.Tensor]):
        """Capture `ops` applied in sequence (each op's outputs feed the next op) into a
        graph.  Returns the list of output lists (static buffers refreshed by replay())."""
        chain_in = [x.contiguous() for x in inputs]
        bufs = []
        for op in ops:
            bufs.append(self.empty_outputs(op))
            self.workspace(op)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):  # warm-up outside the capture
            cur = chain_in
            for op, out in zip(ops, bufs):
                self.run(op, cur, out)
                cur = out
        torch.cuda.current_stream(self.device).wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cur = chain_in
            for op, out in zip(ops, bufs):
                self.run(op, cur, out)
                cur = out
        self._graph = (g, chain_in, bufs)
        return bufs

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph[0].replay()
        return self._graph[2]
