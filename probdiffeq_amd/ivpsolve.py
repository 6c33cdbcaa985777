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
    "dt0",
    "dt0_adaptive",
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


def _set_dt0(cfg, dt0, batch, device):
    """A scalar dt0, or one per instance (what vmap over the reference's solve computes with the
    per-instance proposals of dt0 / dt0_adaptive)."""
    if isinstance(dt0, torch.Tensor) and dt0.numel() > 1:
        if dt0.numel() != batch:
            raise ValueError(f"dt0 has {dt0.numel()} entries for a batch of {batch}")
        cfg.dt0 = float(dt0.reshape(-1)[0])
        cfg.dt0_per_instance = dt0.reshape(-1).to(device=device, dtype=torch.float64).contiguous()
    else:
        cfg.dt0 = float(dt0)


_TIME_DEPENDENT = {"linear"}


def _vf_values(vf, y, t):
    """f(y, t) of a registered vector field, through the device's Taylor-coefficient entry point
    (its coefficient 1 is f(y0, t0)); y (B, d) -> (B, d)."""
    B = y.shape[0]
    params = vf.params_batched(B).to(y.device)
    return capi.taylor_coefficients(vf.name, params, y.contiguous(), float(t), 1)[:, 1, :]


def _initial_values(vf, initial_values):
    if getattr(vf, "is_jet_lifted", False):
        raise ValueError
    u0 = pdq._as_f64(initial_values[0])
    batched = u0.dim() == 2
    return (u0 if batched else u0[None]), batched


def dt0(vf, initial_values, /, scale=0.01, nugget=1e-5, t=0.0):
    """stepsize_initialisers.py:7-21: scale * |u0| / (|f(u0)| + nugget) (2-norms), per instance for
    batched initial values."""
    u0, batched = _initial_values(vf, initial_values)
    f0 = _vf_values(vf, u0, t)
    out = scale * torch.linalg.vector_norm(u0, dim=-1) / (torch.linalg.vector_norm(f0, dim=-1) + nugget)
    return out if batched else out[0]


def dt0_adaptive(vf, initial_values, /, t0, *, error_contraction_rate, rtol, atol):
    """stepsize_initialisers.py:24-67 (Hairer, Norsett, Wanner, Sec. II.4), per instance for batched
    initial values."""
    if len(initial_values) > 1:
        raise ValueError
    y0, batched = _initial_values(vf, initial_values)
    f0 = _vf_values(vf, y0, t0)
    scale = atol + torch.abs(y0) * rtol
    d0, d1 = torch.linalg.vector_norm(y0, dim=-1), torch.linalg.vector_norm(f0, dim=-1)
    h0 = torch.where((d0 < 1e-5) | (d1 < 1e-5), torch.full_like(d0, 1e-6), 0.01 * d0 / d1)
    y1 = y0 + h0[:, None] * f0
    # f at t0 + h0 per instance; of the registry's fields only "linear" depends on t
    if vf.name in _TIME_DEPENDENT:
        f1 = torch.stack([_vf_values(vf, y1[i : i + 1], float(t0) + float(h0[i]))[0] for i in range(y0.shape[0])])
    else:
        f1 = _vf_values(vf, y1, t0)
    d2 = torch.linalg.vector_norm((f1 - f0) / scale, dim=-1) / h0
    h1 = torch.where(
        (d1 <= 1e-15) & (d2 <= 1e-15),
        torch.clamp(h0 * 1e-3, min=1e-6),
        (0.01 / torch.maximum(d1, d2)) ** (1.0 / (error_contraction_rate + 1.0)),
    )
    out = torch.minimum(100.0 * h0, h1)
    return out if batched else out[0]


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
        B = u.tcoeffs.shape[0]
        cfg.rtol, cfg.atol, cfg.eps, cfg.damp = float(rtol), float(atol), float(eps), float(damp)
        _set_dt0(cfg, dt0, B, u.tcoeffs.device)
        cfg.max_attempts = max_attempts
        cfg.lanes_per_instance = lanes_per_instance
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
