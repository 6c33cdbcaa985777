"""Autotuner for the kernel choice: lanes per instance.

For a configuration and a batch (device tensors), every candidate is timed
with HIP events on the current stream: 1 (the lane-per-instance kernel, where
it is instantiated) and 4, 8, 16, 32, 64 (the cooperative kernel with that
group size).  The fastest is recorded in the library's table
(``pdq_set_tuned_lanes``), keyed by configuration, mode and batch-size
bucket, so later calls with ``lanes_per_instance = 0`` use it.  Solves
(``tune_lanes``) and LMLs (``tune_lml_lanes``) have their own entries.  The
reference has no counterpart: XLA chooses its own launch shapes.
"""

from __future__ import annotations

import dataclasses

import torch

from . import capi

CANDIDATES = (1, 4, 8, 16, 32, 64)


def _time(cfg, params, tcoeffs, times, fixed_grid, out, reps):
    capi.solve(cfg, params, tcoeffs, times, fixed_grid=fixed_grid, out=out)  # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        capi.solve(cfg, params, tcoeffs, times, fixed_grid=fixed_grid, out=out)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_lanes(cfg: capi.Config, params, tcoeffs, times, *, fixed_grid=False, candidates=CANDIDATES, reps=2, record=True):
    """Times every candidate on this batch; returns ({lanes: ms or None}, best lanes).

    A candidate the library refuses for this configuration (e.g. 1 for the
    dense factorisation) is reported as None and skipped.
    """
    n_inst, K = times.shape
    out = capi.alloc_solution(cfg, n_inst, K, params.device)
    timings = {}
    for lanes in candidates:
        c = dataclasses.replace(cfg, lanes_per_instance=lanes)
        try:
            timings[lanes] = _time(c, params, tcoeffs, times, fixed_grid, out, reps)
        except capi.PdqError:
            timings[lanes] = None
    valid = {k: v for k, v in timings.items() if v is not None}
    if not valid:
        raise capi.PdqError("no candidate lanes-per-instance value runs this configuration")
    best = min(valid, key=valid.get)
    if record:
        capi.set_tuned_lanes(dataclasses.replace(cfg, lanes_per_instance=0), n_inst, best, mode=1 if fixed_grid else 0)
    return timings, best


LML_MODE = 2  # the table's LML entries (the library's MODE_LML)


def _time_lml(cfg, args, grad, reps):
    capi.lml_value_and_grad(cfg, *args, grad=grad)  # warm-up
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        capi.lml_value_and_grad(cfg, *args, grad=grad)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_lml_lanes(cfg: capi.Config, params, y0, grid, every, data, stds, *, grad=True, candidates=CANDIDATES, reps=2,
                   record=True):
    """Times the LML (with its gradient if `grad`) at every candidate on this batch; returns
    ({lanes: ms or None}, best lanes).

    Candidate 1 is the lane-per-instance LML where it runs this configuration; it is taken by
    every call with lanes_per_instance 0 or 1 and needs no table entry.  The table records the
    fastest cooperative group size, which lanes_per_instance = 0 uses when the lane LML does not
    apply (dense, ts1, solver_mle, other shapes).
    """
    args = (params, y0, grid, every, data, stds)
    timings = {}
    for lanes in candidates:
        c = dataclasses.replace(cfg, lanes_per_instance=lanes)
        try:
            timings[lanes] = _time_lml(c, args, grad, reps)
        except capi.PdqError:
            timings[lanes] = None
    valid = {k: v for k, v in timings.items() if v is not None}
    if not valid:
        raise capi.PdqError("no candidate lanes-per-instance value runs this LML")
    best = min(valid, key=valid.get)
    coop = {k: v for k, v in valid.items() if k != 1}
    if record and coop:
        capi.set_tuned_lanes(dataclasses.replace(cfg, lanes_per_instance=0), params.shape[0], min(coop, key=coop.get),
                             mode=LML_MODE)
    return timings, best
