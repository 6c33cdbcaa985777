"""IVP solver loops (probdiffeq/ivpsolve.py), backed by the gfx950 C ABI.

``solve_adaptive_save_at`` / ``solve_adaptive_terminal_values`` /
``solve_fixed_grid`` keep the reference's factory-then-call shape
(_ivpsolve/solvers_via_adaptive_steps.py:16-148,
_ivpsolve/solvers_via_fixed_steps.py:11-34).  The whole time loop --
attempts, the accept/reject decision, checkpoint interpolation and the
smoother's backward pass -- runs inside one persistent HIP launch per call,
for one instance or for a batch (leading axis of the prior's Taylor
coefficients and of the vector-field parameters).
"""

from __future__ import annotations

import warnings

import torch

from probdiffeq_amd import capi
from probdiffeq_amd import probdiffeq as pdq

__all__ = [
    "control_integral",
    "control_proportional_integral",
    "solve_adaptive_save_at",
    "solve_adaptive_terminal_values",
    "solve_fixed_grid",
]


class control_proportional_integral:
    """controllers.py:24-63."""

    kind = "proportional_integral"

    def __init__(self, *, safety=0.95, factor_min=0.2, factor_max=10.0, exponent_integral=0.3, exponent_proportional=0.4):
        self.safety = safety
        self.factor_min = factor_min
        self.factor_max = factor_max
        self.exponent_integral = exponent_integral
        self.exponent_proportional = exponent_proportional


class control_integral:
    """controllers.py:66-84."""

    kind = "integral"

    def __init__(self, *, safety=0.95, factor_min=0.2, factor_max=10.0):
        self.safety = safety
        self.factor_min = factor_min
        self.factor_max = factor_max


def _times(x, batch, device):
    t = pdq._as_f64(x, device)
    if t.dim() == 1:
        t = t[None].expand(batch, -1)
    if t.shape[0] != batch:
        raise ValueError(f"time grid batch {t.shape[0]} does not match the prior batch {batch}")
    return t.contiguous()


def _raise_on_status(sol_status: torch.Tensor) -> None:
    st = sol_status.detach().cpu()
    if torch.any(st != 0):
        bad = int(torch.nonzero(st)[0, 0])
        name = capi.STATUS.get(int(st[bad]), "unknown")
        msg = f"instance {bad} stopped with status '{name}' ({int((st != 0).sum())} instance(s) failed)"
        raise FloatingPointError(msg)


def solve_adaptive_save_at(*, solver, error, control=None, clip_dt=False, while_loop=None, warn=True, lanes_per_instance=0, check_status=True):
    """solvers_via_adaptive_steps.py:46-148."""
    if not solver.is_suitable_for_save_at and warn:
        msg = f"Solver {solver} should not be used in solve_adaptive_save_at."
        msg += " This is typically caused by the wrong strategy selection."
        msg += " Try using filters or fixed-point smoothers."
        warnings.warn(msg, stacklevel=1)
    if while_loop is not None:
        raise NotImplementedError("the device loop replaces the while_loop injection point")
    if control is None:
        control = control_integral()

    def solve(u, save_at, atol, rtol, dt0=0.1, eps=1e-8, damp=0.0, max_attempts=0):
        cfg = pdq.make_config(u, solver, error, control, clip_dt)
        cfg.rtol, cfg.atol, cfg.dt0, cfg.eps, cfg.damp = float(rtol), float(atol), float(dt0), float(eps), float(damp)
        cfg.max_attempts = max_attempts
        cfg.lanes_per_instance = lanes_per_instance
        B = u.tcoeffs.shape[0]
        params = solver.constraint.ode.params_batched(B)
        times = _times(save_at, B, u.tcoeffs.device)
        out = capi.solve(cfg, params, u.tcoeffs, times)
        sol = pdq.build_solution(cfg, out, batched=u.batched)
        if check_status:
            _raise_on_status(out["status"])
        return sol

    return solve


def solve_adaptive_terminal_values(*, solver, error, control=None, clip_dt=True, while_loop=None, lanes_per_instance=0):
    """solvers_via_adaptive_steps.py:16-43."""
    solve_save_at = solve_adaptive_save_at(solver=solver, error=error, control=control, clip_dt=clip_dt, warn=False, while_loop=while_loop, lanes_per_instance=lanes_per_instance)

    def solve(u, /, *, t0, t1, atol, rtol, dt0=0.1, eps=1e-8, damp=0.0):
        sol = solve_save_at(u, torch.tensor([float(t0), float(t1)], dtype=torch.float64), atol=atol, rtol=rtol, dt0=dt0, eps=eps, damp=damp)
        return _last(sol, u.batched)

    return solve


def _last(sol, batched):
    import dataclasses

    def take(x):
        if isinstance(x, torch.Tensor):
            return x[:, -1] if batched and x.dim() >= 2 else (x[-1] if x.dim() >= 1 else x)
        if dataclasses.is_dataclass(x):
            return type(x)(**{f.name: take(getattr(x, f.name)) for f in dataclasses.fields(x)})
        return x

    return take(sol)


def solve_fixed_grid(*, solver, lanes_per_instance=0, check_status=True):
    """solvers_via_fixed_steps.py:11-34."""
    if not solver.is_suitable_for_save_every_step:
        msg = f"Solver {solver} should not be used in solve_adaptive_save_every_step/solve_fixed_grid."
        msg += " This is typically caused by using a fixed-point smoother."
        msg += " Try using filters or fixed-interval smoothers instead."
        warnings.warn(msg, stacklevel=1)

    def solve(u, /, *, grid, damp: float = 0.0):
        cfg = pdq.make_config(u, solver)
        cfg.damp = float(damp)
        cfg.lanes_per_instance = lanes_per_instance
        B = u.tcoeffs.shape[0]
        params = solver.constraint.ode.params_batched(B)
        times = _times(grid, B, u.tcoeffs.device)
        out = capi.solve(cfg, params, u.tcoeffs, times, fixed_grid=True)
        sol = pdq.build_solution(cfg, out, batched=u.batched)
        if check_status:
            _raise_on_status(out["status"])
        return sol

    return solve
