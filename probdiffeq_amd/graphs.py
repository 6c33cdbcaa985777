"""Captured launch graphs of a batched solve.

A captured solve allocates nothing and clears nothing with memset nodes:
* save_at solves go through pdq_solve_adaptive_save_at_noalloc, whose IWP
  matrices and work-queue counter live in the library's fixed device
  workspace and are written by small kernels, so every kernel (lane,
  cooperative, blockdiag) can be captured;
* fixed-grid solves have no allocation-free entry point and are captured on
  the lane-per-instance kernel only (one launch, its working state in LDS,
  every output element written).
The graph reads its inputs from, and writes its outputs to, the same device
buffers on every replay: update the inputs in place between replays.  Replays
on one GPU share the fixed workspace with every other no-allocation call on
that GPU: do not overlap them.
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
        if fixed_grid:
            cfg = dataclasses.replace(cfg, lanes_per_instance=1 if cfg.lanes_per_instance == 0 else cfg.lanes_per_instance)
            if cfg.lanes_per_instance != 1 or cfg.factorisation != "isotropic":
                msg = "fixed-grid graphs need the lane-per-instance kernel (isotropic, lanes_per_instance 0 or 1);"
                msg += " save_at solves capture every kernel"
                raise NotImplementedError(msg)
        self.cfg, self.params, self.tcoeffs, self.times, self.fixed_grid = cfg, params, tcoeffs, times, fixed_grid
        n_inst, K = times.shape
        self.out = capi.alloc_solution(cfg, n_inst, K, times.device)
        no_alloc = not fixed_grid
        # warm-up outside the capture: loads the code objects (and, for fixed grids, caches the
        # IWP matrices)
        capi.solve(cfg, params, tcoeffs, times, fixed_grid=fixed_grid, out=self.out, no_alloc=no_alloc)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            capi.solve(cfg, params, tcoeffs, times, fixed_grid=fixed_grid, out=self.out, no_alloc=no_alloc)
        torch.cuda.synchronize()

    def replay(self) -> dict:
        self.graph.replay()
        return self.out
