"""torch.autograd wrappers of the C ABI.

``LML`` is the differentiable marginal log-likelihood of the parameter
estimation workflow (examples/E2_parameter_estimation_optax.py): forward
runs the fixed-point-smoother solve and ``evaluate_lml`` on the device and
obtains the gradient with respect to the vector-field parameters and the
initial value in the same launch; backward scales it by the incoming
gradient.  There is no CPU path.
"""

from __future__ import annotations

import dataclasses
from typing import Any

import torch

from probdiffeq_amd import capi


@dataclasses.dataclass
class LMLProblem:
    """The solve behind the likelihood (mirrors the example's solver setup).

    ``grid``: fixed time grid (K_grid,); ``every``: observation stride on the
    grid (observations at grid[::every]); the solver is ``solver`` /
    ``solver_dynamic`` with ``strategy_smoother_fixedpoint`` and ts0.
    """

    vector_field: str
    num_derivatives: int
    ode_dim: int
    grid: torch.Tensor
    every: int
    factorisation: str = "isotropic"
    solver: str = "solver"
    base_output_scale: float = 1.0
    init_std: float = 0.0
    lanes_per_instance: int = 0
    # the gradient: "forward" (tangents) or "reverse" (the adjoint with checkpoints)
    lml_gradient: str = "forward"
    # "iwp" or "exponential" (dense): prior_drift is the drift's bottom block, a device tensor
    # (d, (num_derivatives + 1) d), as in capi.Config
    prior: str = "iwp"
    prior_drift: Any = None

    def config(self, n_params: int) -> capi.Config:
        return capi.Config(
            factorisation=self.factorisation,
            num_derivatives=self.num_derivatives,
            ode_dim=self.ode_dim,
            strategy="fixedpoint",
            solver=self.solver,
            vector_field=self.vector_field,
            n_params=n_params,
            base_output_scale=self.base_output_scale,
            init_std=self.init_std,
            lanes_per_instance=self.lanes_per_instance,
            lml_gradient=self.lml_gradient,
            prior=self.prior,
            prior_drift=self.prior_drift,
        )


class LML(torch.autograd.Function):
    """value = loss_lml_timeseries(data, posterior=solve(params, y0), std)."""

    @staticmethod
    def forward(ctx, params, y0, data, stds, problem: LMLProblem, average: bool, tcoeff_index: int = 0,
                solve: str = "lstsq_svd"):
        cfg = problem.config(params.shape[-1])
        cfg.lml_tcoeff_index = int(tcoeff_index)
        cfg.lml_solve = solve
        need_grad = params.requires_grad or y0.requires_grad
        value, grad = capi.lml_value_and_grad(
            cfg,
            params.detach().contiguous(),
            y0.detach().contiguous(),
            problem.grid.contiguous(),
            problem.every,
            data.contiguous(),
            stds.contiguous(),
            average=average,
            grad=need_grad,
        )
        ctx.n_params = params.shape[-1]
        ctx.save_for_backward(grad if grad is not None else value)
        ctx.has_grad = grad is not None
        return value

    @staticmethod
    def backward(ctx, g):
        if not ctx.has_grad:
            return None, None, None, None, None, None, None, None
        (grad,) = ctx.saved_tensors
        full = grad * g[:, None]
        return full[:, : ctx.n_params], full[:, ctx.n_params :], None, None, None, None, None, None


def lml(problem: LMLProblem, data, std, *, params, y0, average_pdfs=True, tcoeff_index=0, solve="lstsq_svd"):
    """Batched LML (one value per instance) with autograd support; the data observe Taylor
    coefficient ``tcoeff_index``; ``solve`` ("lstsq_svd" or "solve_triu") is the gains' solve."""
    return LML.apply(params, y0, data, std, problem, average_pdfs, tcoeff_index, solve)
