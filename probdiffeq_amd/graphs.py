"""Captured launch graphs of a batched solve.

A solve on the lane-per-instance kernel is one kernel launch that allocates
nothing and clears nothing (its working state is LDS; every output element is
written), so it can be captured into a HIP graph once and replayed with no
host-side launch work.  The graph reads its inputs from, and writes its
outputs to, the same device buffers on every replay: update the inputs in
place between replays.  The cooperative kernel allocates its work-queue
counter per call and is not captured.
"""

from __future__ import annotations

import dataclasses

import torch

from . import capi


class SolveGraph:
    """One batched solve captured as a HIP graph (torch.cuda.CUDAGraph).

    params (B, P), tcoeffs (B, N, d) and times (B, K) are device tensors the
    graph keeps reading; ``replay()`` runs the solve and returns the output
    dict (the same tensors every time).
    """

    def __init__(self, cfg: capi.Config, params, tcoeffs, times, *, fixed_grid: bool = False):
        cfg = dataclasses.replace(cfg, lanes_per_instance=1 if cfg.lanes_per_instance == 0 else cfg.lanes_per_instance)
        if cfg.lanes_per_instance != 1:
            raise NotImplementedError("graph capture needs the lane-per-instance kernel (lanes_per_instance 0 or 1)")
        self.cfg, self.params, self.tcoeffs, self.times, self.fixed_grid = cfg, params, tcoeffs, times, fixed_grid
        n_inst, K = times.shape
        self.out = capi.alloc_solution(cfg, n_inst, K, times.device)
        # warm-up outside the capture: loads the code object and caches the IWP matrices
        capi.solve(cfg, params, tcoeffs, times, fixed_grid=fixed_grid, out=self.out)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            capi.solve(cfg, params, tcoeffs, times, fixed_grid=fixed_grid, out=self.out)
        torch.cuda.synchronize()

    def replay(self) -> dict:
        self.graph.replay()
        return self.out
