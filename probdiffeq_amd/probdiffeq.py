"""Strategy/solver API of probdiffeq, backed by the gfx950 C ABI.

Mirrors the names, arguments and error behaviour of the reference's
``probdiffeq.probdiffeq`` facade (probdiffeq/probdiffeq.py:6-17) for the
configurations the HIP kernels implement.  Configuration objects only
collect parameters; every computation happens in one batched launch in
``probdiffeq_amd.ivpsolve`` (or ``loss_lml_timeseries``).  A configuration
the kernels do not implement raises ``NotImplementedError``; nothing falls
back to the CPU.

Vector fields: device code cannot call Python callables, so ODEs come from
the registry in this module (``lotka_volterra``, ``van_der_pol``,
``burgers``, ``linear``, ``logistic``), each a ``JetOde`` with
``num_tcoeffs_in_args=1`` and ``tcoeff_indices_output=[1]`` like
``problems.ode`` (problems.py:283-295).  Parameters may carry a leading
batch dimension; a batch is solved in one launch (the reference uses
jax.vmap).
"""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np
import torch

from probdiffeq_amd import capi

__all__ = [
    "JetOde",
    "MarkovSequence",
    "Normal",
    "ProbabilisticSolution",
    "SmoothingSolution",
    "burgers",
    "error_norm_rms_then_scale",
    "error_norm_scale_then_rms",
    "error_residual_std",
    "error_state_std",
    "jacobian_materialize",
    "jacobian_monte_carlo_fwd",
    "jacobian_monte_carlo_rev",
    "JetOdeAutonomous",
    "jetexpand_ode_doubling_unroll",
    "jetexpand_ode_padded_scan",
    "jetexpand_ode_unroll",
    "jetexpand_ode_via_jvp",
    "preconditioner_taylor",
    "system_matrices_1d_iwp",
    "taylor_point_prior",
    "ode_autonomous_order_arbitrary",
    "linear",
    "logistic",
    "loss_lml_terminal_values",
    "loss_lml_timeseries",
    "lotka_volterra",
    "ode",
    "solver",
    "solver_dynamic",
    "solver_mle",
    "state_space_model_blockdiag",
    "state_space_model_dense",
    "state_space_model_isotropic",
    "strategy_filter",
    "strategy_smoother_fixedinterval",
    "strategy_smoother_fixedpoint",
    "van_der_pol",
]


def _device():
    return capi.require_device()


def _as_f64(x, device=None):
    if isinstance(x, torch.Tensor):
        return x.to(dtype=torch.float64, device=device or _device()).contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=device or _device()).contiguous()


# ----------------------------------------------------------------- problems
@dataclasses.dataclass
class JetOde:
    """A registered vector field u' = f(u, t; params) (problems.py:170-210)."""

    name: str
    params: Any  # (P,) or (B, P)
    ode_dim: int
    num_tcoeffs_in_args: int = 1
    tcoeff_indices_output: tuple = (1,)
    jacobian: Any = None

    def params_batched(self, batch: int | None = None) -> torch.Tensor:
        p = _as_f64(self.params)
        if p.dim() == 1:
            p = p[None]
        if batch is not None and p.shape[0] == 1 and batch > 1:
            p = p.expand(batch, -1).contiguous()
        return p


def lotka_volterra(a=0.5, b=0.05, c=0.5, d=0.05) -> JetOde:
    """dy1 = a y1 - b y1 y2, dy2 = -c y2 + d y1 y2 (backend/ode.py:30-51)."""
    p = torch.stack([_as_f64(v) for v in (a, b, c, d)], dim=-1)
    return JetOde("lotka_volterra", p, 2)


def van_der_pol(mu=1.0) -> JetOde:
    """y1' = y2, y2' = mu ((1 - y1^2) y2) - y1."""
    p = _as_f64(mu)
    return JetOde("van_der_pol", p[..., None] if p.dim() <= 1 else p, 2)


def burgers(num_points: int, viscosity=0.01, length=1.0) -> JetOde:
    """u_t + u u_x = nu u_xx on a periodic grid, central differences."""
    return JetOde("burgers", _as_f64([viscosity, length / num_points]), num_points)


def linear(lam) -> JetOde:
    """u' = lam * u + t (the reference's calibration-equivalence test ODE)."""
    lam = _as_f64(lam)
    return JetOde("linear", lam, lam.shape[-1])


def logistic(r=2.0, ode_dim=1) -> JetOde:
    """u' = r u (1 - u) (examples/A0_quickstart.py)."""
    p = _as_f64(r)
    return JetOde("logistic", p[..., None] if p.dim() <= 1 else p, ode_dim)


def ode(func, /, *, jacobian=None):
    """problems.ode (problems.py:283-295): only registered vector fields run on the device."""
    if isinstance(func, JetOde):
        if jacobian is not None:
            func = dataclasses.replace(func, jacobian=jacobian)
        return func
    msg = "Arbitrary Python vector fields cannot run inside the gfx950 kernels; use a registered"
    msg += " vector field (lotka_volterra, van_der_pol, burgers, linear, logistic)."
    raise NotImplementedError(msg)


class jacobian_materialize:
    """jacobians.py:78-113; the registry provides analytic Jacobians, so traces are exact."""

    def __init__(self, *, jacfun=None) -> None:
        self.jacfun = jacfun

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}(jacfun={self.jacfun})"


class _JacobianMonteCarlo:
    """An exact dense Jacobian (dense ts1); Hutchinson estimates of the isotropic trace and the
    blockdiag diagonal from `num_probes` Rademacher probes (1..64).  The kernels draw one probe set
    per linearisation from a counter-based stream keyed by `seed` and the accepted steps (a
    rejected attempt's retry reuses its probes, as the reference's key travels with the accepted
    solution); it is not JAX's threefry stream, so single estimates differ from the reference's
    while following the same estimator."""

    mode = ""

    def __init__(self, *, seed=1, num_probes=10) -> None:
        self.seed = seed
        self.num_probes = num_probes

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}(seed={self.seed}, num_probes={self.num_probes})"


class jacobian_monte_carlo_fwd(_JacobianMonteCarlo):
    """jacobians.py:116-192 (forward mode): diagonal entry k is the probe mean of (J v)_k v_k."""

    mode = "monte_carlo_fwd"


class jacobian_monte_carlo_rev(_JacobianMonteCarlo):
    """jacobians.py:195-271 (reverse mode; the reference's default handler of ode(),
    problems.py:283-295): diagonal entry k is the probe mean of (v^T J)_k v_k."""

    mode = "monte_carlo_rev"


def _jacobian_handler(jac):
    """(pdq_config.jacobian, probes, seed) of a Jacobian handler; None is ode()'s default,
    jacobian_monte_carlo_rev(seed=1, num_probes=10) (problems.py:283-295)."""
    if jac is None:
        jac = jacobian_monte_carlo_rev()
    if isinstance(jac, jacobian_materialize):
        return "exact", 10, 1
    if isinstance(jac, _JacobianMonteCarlo):
        probes = int(jac.num_probes)
        if not 1 <= probes <= 64:
            raise NotImplementedError(f"the kernels draw 1 to 64 probes per linearisation (got num_probes={probes})")
        return jac.mode, probes, int(jac.seed) & 0xFFFFFFFF
    raise TypeError(f"unknown Jacobian handler {jac!r}")


class JetOdeAutonomous:
    """problems.JetOdeAutonomous (problems.py:254-281): an autonomous u^(k) = f(u, u', ...).

    Only the exponential priors take one: its drift is a torch-traceable Python function of the
    Taylor coefficients, differentiated on the host when the prior is built (it never runs in a
    kernel)."""

    def __init__(self, autonomous, *, num_tcoeffs_in_args, jacobian=None, tcoeff_indices_output=None):
        self.autonomous = autonomous
        self.num_tcoeffs_in_args = num_tcoeffs_in_args
        self.jacobian = jacobian
        self.tcoeff_indices_output = [num_tcoeffs_in_args] if tcoeff_indices_output is None else tcoeff_indices_output


def ode_autonomous_order_arbitrary(func, /, *, num_tcoeffs_in_args, jacobian=None):
    """problems.ode_autonomous_order_arbitrary (problems.py:385-401): func(u, du, ..., /) -> f."""
    return JetOdeAutonomous(lambda *, jet_coords: func(*jet_coords[:num_tcoeffs_in_args]), num_tcoeffs_in_args=num_tcoeffs_in_args,
                            jacobian=jacobian)


def _ioup_ode(linop, N):
    """prior_ornstein_uhlenbeck_integrated's drift (ssm_impl_api.py:381-410): f = linop(u^(N-1))."""
    return JetOdeAutonomous(lambda *, jet_coords: [linop(jet_coords[-1])], num_tcoeffs_in_args=N)


def _matern_ode(length_scale, N):
    """prior_matern's drift (ssm_impl_api.py:443-477): f = -sum_i C(D, i) z^(D-i) u^(i) with
    z = sqrt(2 (D - 1/2)) / length_scale and D = N Taylor coefficients."""
    from math import comb, sqrt

    def autonomous(*, jet_coords):
        D = len(jet_coords)
        z = sqrt(2 * (D - 0.5)) / length_scale
        return [-sum(comb(D, i) * z ** (D - i) * c for i, c in enumerate(jet_coords))]

    return JetOdeAutonomous(autonomous, num_tcoeffs_in_args=N)


def _linear_drift(ode, tc):
    """The Jacobian of the flattened f (d,) with respect to the flattened Taylor coefficients (N d,),
    derivative-major (ssm_impl_dense.py:587-596), at the coefficients tc (N, d)."""
    N, d = tc.shape

    def vf_flat(x):
        coeffs = [x[i * d : (i + 1) * d] for i in range(N)]
        out = ode.autonomous(jet_coords=coeffs)
        if isinstance(out, (list, tuple)):
            out = torch.cat([torch.as_tensor(o, dtype=torch.float64).reshape(-1) for o in out])
        return out.reshape(-1)

    J = torch.autograd.functional.jacobian(vf_flat, tc.reshape(-1).to(torch.float64))
    if J.shape != (d, N * d):
        raise ValueError(f"the prior's drift maps (N d,) = ({N * d},) to shape {tuple(J.shape)}, expected ({d}, {N * d})")
    return J.contiguous()


def jetexpand_ode_padded_scan(*, num: int):
    """Taylor coefficients by Taylor-mode recursion (jet_expansion_algorithms.py:49-103)."""

    def expand(vf: JetOde, inits, /, *, t):
        if not isinstance(vf, JetOde):
            raise TypeError(vf)
        (u0,) = inits
        u0 = _as_f64(u0)
        y0 = u0[None] if u0.dim() == 1 else u0
        params = vf.params_batched(y0.shape[0])
        if params.shape[0] != y0.shape[0]:
            raise ValueError(f"batch sizes differ: params {params.shape[0]}, inits {y0.shape[0]}")
        tc = capi.taylor_coefficients(vf.name, params, y0.contiguous(), float(t), num)
        coeffs = [tc[:, i, :] for i in range(num + 1)]
        if u0.dim() == 1:
            coeffs = [c[0] for c in coeffs]
        return coeffs, {}

    return expand


def jetexpand_ode_unroll(*, num: int):
    """jet_expansion_algorithms.py:110-152.  The same Taylor coefficients as
    jetexpand_ode_padded_scan; unrolling versus scanning is a JAX compilation choice, and the device
    runs the same recursion."""
    return jetexpand_ode_padded_scan(num=num)


def jetexpand_ode_via_jvp(*, num: int):
    """jet_expansion_algorithms.py:180-212 (recursive forward mode): the same coefficients, by the
    device's Taylor-mode recursion."""
    return jetexpand_ode_padded_scan(num=num)


def jetexpand_ode_doubling_unroll(*, num_doublings: int):
    """jet_expansion_algorithms.py:228-262 (Newton's doubling): num_doublings doublings give
    2^(num_doublings + 1) - 1 Taylor coefficients; the device's recursion computes the same ones."""
    return jetexpand_ode_padded_scan(num=2 ** (num_doublings + 1) - 2)


def system_matrices_1d_iwp(num_derivatives: int):
    """utilities.py:57-71: the latent IWP transition A and noise factor Q ((nu+1) x (nu+1)), as the
    library computes them (host only)."""
    A, Q = capi.system_matrices_1d_iwp(num_derivatives)
    return torch.as_tensor(A), torch.as_tensor(Q)


def preconditioner_taylor(num_derivatives: int):
    """utilities.py:74-84: dt -> (p, p^-1) with p_i = dt^(nu-i) / (nu-i)!."""
    from math import factorial

    powers = torch.arange(num_derivatives, -1, -1, dtype=torch.float64)
    scales = torch.tensor([float(factorial(int(k))) for k in powers], dtype=torch.float64)

    def precon(dt):
        dt = torch.as_tensor(dt, dtype=torch.float64)
        return dt**powers / scales, dt ** (-powers) * scales

    return precon


class taylor_point_prior:
    """taylor_points.py:150-156: linearise ts1 at the prior (predicted) mean, the device path's
    (and the reference's) default Taylor point."""

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}()"


# ------------------------------------------------------------ distributions
@dataclasses.dataclass
class Normal:
    """Flattened normal of one factorisation (ssm_impl_*.py Normal classes).

    ``mean_flat`` / ``cholesky_flat`` carry leading (batch, time) axes:
    dense (.., n) / (.., n, n); isotropic (.., N, d) / (.., N, N);
    blockdiag (.., d, N) / (.., d, N, N).
    """

    mean_flat: torch.Tensor
    cholesky_flat: torch.Tensor
    factorisation: str
    num_coeffs: int
    ode_dim: int

    @property
    def mean(self) -> list:
        """List of Taylor coefficients, each (.., d) (tree_flatten.unflatten_array)."""
        N, d = self.num_coeffs, self.ode_dim
        if self.factorisation == "dense":
            return [self.mean_flat[..., i * d : (i + 1) * d] for i in range(N)]
        if self.factorisation == "isotropic":
            return [self.mean_flat[..., i, :] for i in range(N)]
        return [self.mean_flat[..., :, i] for i in range(N)]

    @property
    def std(self) -> list:
        """Row norms of the Cholesky factor (ssm_impl_dense.py:141-150 & co.)."""
        rn = torch.linalg.vector_norm(self.cholesky_flat, dim=-1)
        N, d = self.num_coeffs, self.ode_dim
        if self.factorisation == "dense":
            return [rn[..., i * d : (i + 1) * d] for i in range(N)]
        if self.factorisation == "isotropic":
            return [rn[..., i] for i in range(N)]
        return [rn[..., :, i] for i in range(N)]

    def to_multivariate_normal(self):
        """(mean, covariance) in the reference's flattened coordinates."""
        L = self.cholesky_flat
        cov = L @ L.transpose(-1, -2)
        if self.factorisation == "dense":
            return self.mean_flat, cov
        if self.factorisation == "isotropic":
            d = self.ode_dim
            eye = torch.eye(d, dtype=cov.dtype, device=cov.device)
            full = torch.einsum("...nm,dt->...ndmt", cov, eye)
            n = self.mean_flat.shape[-2] * d
            return self.mean_flat.reshape(*self.mean_flat.shape[:-2], n), full.reshape(*full.shape[:-4], n, n)
        raise NotImplementedError("blockdiag to_multivariate_normal")


@dataclasses.dataclass
class LatentConditional:
    """Backward conditionals p(x_k | x_{k+1}) in latent coordinates (ssm_impl_api.py:35-132)."""

    A: torch.Tensor
    noise: Normal
    to_latent: torch.Tensor
    to_observed: torch.Tensor


@dataclasses.dataclass
class MarkovSequence:
    """estimators_and_losses.py:121-273: a terminal marginal and backward conditionals."""

    marginal: Normal
    conditional: LatentConditional
    reverse: bool = True

    def sample(self, key, *, shape: tuple = ()):
        """estimators_and_losses.py:233-273 on the device (dense and isotropic).  ``key`` is an int seed
        or a ``torch.Generator`` on the device (torch has no JAX keys); the standard normal draws come
        from it, so samples match the reference's law, not its bits.  Returns (*shape, K, N, d)
        (isotropic) or (*shape, K, N d) (dense, the derivative-major mean layout), with a leading
        instance axis for a batched solution."""
        gen = key
        if not isinstance(key, torch.Generator):
            gen = torch.Generator(device=self.marginal.mean_flat.device)
            gen.manual_seed(int(key))
        shape = tuple(shape)
        n = int(np.prod(shape)) if shape else 1
        dense = self.marginal.factorisation == "dense"
        batched = self.marginal.mean_flat.dim() == (2 if dense else 3)
        B = self.marginal.mean_flat.shape[0] if batched else 1
        K = self.conditional.A.shape[-3] + 1
        N = self.marginal.num_coeffs
        draws = N * self.marginal.ode_dim if dense else N
        base = torch.randn(B, n, K, draws, generator=gen, dtype=torch.float64, device=self.marginal.mean_flat.device)
        out = self.sample_from_draws(base if batched else base[0])
        lead = ((B,) if batched else ()) + shape
        tail = out.shape[-2:] if dense else out.shape[-3:]
        return out.reshape(*lead, *tail)

    def sample_from_draws(self, base):
        """The sampler with given standard normal draws base (S, K, n), or (B, S, K, n) for a batched
        solution, n the covariance dimension (N isotropic, N d dense): row K - 1 of each sample draws
        the terminal marginal, row k the conditional k."""
        m = self.marginal
        if m.factorisation not in ("isotropic", "dense"):
            raise NotImplementedError("sampling is implemented for the dense and isotropic factorisations")
        dense = m.factorisation == "dense"
        N, d = m.num_coeffs, m.ode_dim
        n = N * d if dense else N
        batched = m.mean_flat.dim() == (2 if dense else 3)
        b = base if batched else base[None]
        c = self.conditional
        B = b.shape[0]
        K = c.A.shape[-3] + 1
        out = {
            "term_mean": (m.mean_flat if batched else m.mean_flat[None]).reshape(B, N * d).contiguous(),
            "term_chol": (m.cholesky_flat if batched else m.cholesky_flat[None]).reshape(B, n * n).contiguous(),
            "cond_A": (c.A if batched else c.A[None]).reshape(B, K - 1, n * n).contiguous(),
            "cond_mean": (c.noise.mean_flat if batched else c.noise.mean_flat[None]).reshape(B, K - 1, N * d).contiguous(),
            "cond_chol": (c.noise.cholesky_flat if batched else c.noise.cholesky_flat[None]).reshape(B, K - 1, n * n).contiguous(),
            "cond_tl": (c.to_latent if batched else c.to_latent[None]).reshape(B, K - 1, n).contiguous(),
            "cond_to": (c.to_observed if batched else c.to_observed[None]).reshape(B, K - 1, n).contiguous(),
        }
        cfg = capi.Config(factorisation=m.factorisation, num_derivatives=N - 1, ode_dim=d, strategy="fixedpoint")
        samples = capi.sample(cfg, out, K, b.contiguous())
        samples = samples.reshape(B, b.shape[1], K, N * d) if dense else samples.reshape(B, b.shape[1], K, N, d)
        return samples if batched else samples[0]


@dataclasses.dataclass
class SmoothingSolution:
    posterior: MarkovSequence
    filtering: Normal


@dataclasses.dataclass
class ProbabilisticSolution:
    """solvers.py:33-69, batched: every field has a leading instance axis."""

    t: torch.Tensor
    u: Normal
    solution_full: Any
    output_scale: torch.Tensor
    num_steps: torch.Tensor
    attempts: torch.Tensor  # attempted steps (accepted + rejected), per instance
    status: torch.Tensor    # 0 ok, 1 max attempts, 2 non-finite, 3 dt underflow
    config: Any = None      # the capi.Config of the solve (the reference keeps solution.prior)

    def squeeze(self):
        """Drop the instance axis of a single-instance solve."""

        def sq(x):
            if isinstance(x, torch.Tensor):
                return x[0]
            if dataclasses.is_dataclass(x):
                return type(x)(**{f.name: sq(getattr(x, f.name)) for f in dataclasses.fields(x)})
            return x

        return sq(self)


# ------------------------------------------------------- state-space models
@dataclasses.dataclass
class Prior:
    """An integrated Wiener process prior (prior_wiener_integrated), or (kind "exponential",
    dense) an exponential-integrator prior whose drift's last d rows are `drift` (d, N d)."""

    factorisation: str
    tcoeffs: torch.Tensor  # (B, N, d)
    init_std: float
    base_output_scale: float
    batched: bool
    kind: str = "iwp"
    drift: torch.Tensor | None = None
    init_std_rows: torch.Tensor | None = None  # a std per covariance row, or None (init_std)

    @property
    def num_derivatives(self) -> int:
        return self.tcoeffs.shape[1] - 1

    @property
    def ode_dim(self) -> int:
        return self.tcoeffs.shape[2]


@dataclasses.dataclass
class Constraint:
    kind: str  # "ts0" | "ts1"
    ode: JetOde
    residual_order: int = 2


class _StateSpaceModel:
    factorisation = ""

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}()"

    def prior_wiener_integrated(self, tcoeffs, /, *, is_exact=True, inexact_eps=1e-6, diffuse_derivatives=0, diffuse_eps=1.0, output_scale=None):
        """ssm_impl_*.py prior_wiener_integrated (dense :458-514, iso :409-482, bd :463-547)."""
        if isinstance(tcoeffs, torch.Tensor) or isinstance(tcoeffs, np.ndarray):
            msg = "Mean must be a pytree, not an array."
            raise TypeError(msg)
        coeffs = [_as_f64(c) for c in tcoeffs]
        shape0 = coeffs[0].shape
        for i, c in enumerate(coeffs):
            if c.shape != shape0:
                msg = f"All leaves of the mean must have the same shape. However, leaf {i} has shape {tuple(c.shape)}, while leaf 0 has shape {tuple(shape0)}"
                raise ValueError(msg)
        rows = None
        if diffuse_derivatives or not isinstance(is_exact, bool):
            rows = self._init_std_rows(coeffs, is_exact, inexact_eps, int(diffuse_derivatives), float(diffuse_eps))
            coeffs = coeffs + [torch.zeros_like(coeffs[0]) for _ in range(int(diffuse_derivatives))]
            is_exact = True
        batched = coeffs[0].dim() == 2
        tc = torch.stack(coeffs, dim=-2)
        if not batched:
            tc = tc[None]
        init_std = 0.0 if is_exact else float(inexact_eps)
        base = 1.0
        if output_scale is not None:
            s = np.asarray(output_scale.cpu() if isinstance(output_scale, torch.Tensor) else output_scale, dtype=np.float64)
            expect = () if self.factorisation == "isotropic" else (tc.shape[-1],)
            if s.shape not in [expect, ()]:
                msg = f"The base-scale has the wrong shape. Expected: {expect}. Received: {s.shape}."
                raise ValueError(msg)
            if s.size > 1 and not np.all(s == s.flat[0]):
                raise NotImplementedError("non-uniform base output scales are not implemented on the device path")
            base = float(s.flat[0])
        return Prior(self.factorisation, tc.contiguous(), init_std, base, batched, init_std_rows=rows)

    def _init_std_rows(self, coeffs, is_exact, inexact_eps, diffuse, diffuse_eps):
        """_tcoeffs_standard_deviation and _add_diffuse_derivatives (ssm_impl_isotropic.py:519-567,
        dense alike) as one std per covariance row: 0 for exact coefficients, inexact_eps otherwise,
        diffuse_eps for the appended zero-mean derivatives.  Isotropic stds are one scalar per
        derivative; dense ones one per element, derivative-major."""
        if self.factorisation not in ("dense", "isotropic"):
            raise NotImplementedError("element-wise is_exact and diffuse derivatives need the dense or isotropic factorisation")
        d = coeffs[0].shape[-1]
        width = d if self.factorisation == "dense" else 1
        flags = [is_exact] * len(coeffs) if isinstance(is_exact, bool) else list(is_exact)
        if len(flags) != len(coeffs):
            raise ValueError(f"is_exact has {len(flags)} leaves; the mean has {len(coeffs)}")
        rows = []
        for e in flags:
            e = np.asarray(e.cpu() if isinstance(e, torch.Tensor) else e, dtype=bool)
            if self.factorisation == "isotropic" and e.ndim != 0:
                raise ValueError("isotropic models take one is_exact flag per Taylor coefficient")
            rows.append(np.broadcast_to(np.where(e, 0.0, float(inexact_eps)), (width,)))
        rows += [np.full(width, diffuse_eps)] * diffuse
        return torch.tensor(np.concatenate(rows), dtype=torch.float64, device=coeffs[0].device)

    def prior_exponential(self, ode, tcoeffs, /, *, is_exact=True, inexact_eps=1e-6, diffuse_derivatives=0, diffuse_eps=1.0, output_scale=None):
        """ssm_impl_dense.py:516-629: the prior of the linear SDE whose drift is the IWP's shift with
        the Jacobian of the linear autonomous `ode` (jet_coords -> f, one argument per Taylor
        coefficient) as its last block row; every transition is exp_gram_cholesky (the device computes
        it per step).  The Jacobian is taken on the host (FP64 autograd) at the Taylor coefficients and
        is shared by the batch, so it must be the same for every instance (a linear drift)."""
        if self.factorisation != "dense":
            name = {"isotropic": "Isotropic", "blockdiag": "Block-diagonal"}[self.factorisation]
            msg = f"{name} exponential priors have not been implemented (yet.). If you need them, reach out."
            raise NotImplementedError(msg)
        base = self.prior_wiener_integrated(tcoeffs, is_exact=is_exact, inexact_eps=inexact_eps, diffuse_derivatives=diffuse_derivatives,
                                            diffuse_eps=diffuse_eps, output_scale=output_scale)
        # the order is checked against the coefficients after the diffuse padding
        # (ssm_impl_dense.py:553-573)
        n_coeffs = base.tcoeffs.shape[-2]
        if ode.num_tcoeffs_in_args != n_coeffs:
            msg = "The exponential prior does not match the Taylor coefficients in the SSM."
            msg += f" The passed ODE has order {ode.num_tcoeffs_in_args}, whereas the state-space model"
            msg += f" includes {n_coeffs} Taylor coefficients."
            raise TypeError(msg)
        tc = base.tcoeffs.detach().to("cpu")
        drifts = [_linear_drift(ode, tc[i]) for i in sorted({0, tc.shape[0] - 1})]
        if not torch.equal(drifts[0], drifts[-1]):
            raise ValueError("the exponential prior's drift must be linear (the same Jacobian for every instance)")
        return dataclasses.replace(base, kind="exponential", drift=drifts[0].to(base.tcoeffs.device).contiguous())

    def prior_ornstein_uhlenbeck_integrated(self, linop, tcoeffs, /, **kwargs):
        """ssm_impl_api.py:381-410: the exponential prior with f(u, ..., u^(nu)) = linop(u^(nu))."""
        N = len(tcoeffs) + kwargs.get("diffuse_derivatives", 0)
        return self.prior_exponential(_ioup_ode(linop, N), tcoeffs, **kwargs)

    def prior_matern(self, length_scale, tcoeffs, /, **kwargs):
        """ssm_impl_api.py:443-477: the exponential prior of the Matern drift (_matern_ode)."""
        N = len(tcoeffs) + kwargs.get("diffuse_derivatives", 0)
        return self.prior_exponential(_matern_ode(length_scale, N), tcoeffs, **kwargs)

    def constraint_ode_ts0(self, ode, /):
        """ssm_impl_api.py:520-534."""
        if not isinstance(ode, JetOde):
            raise TypeError(ode)
        return Constraint("ts0", ode)

    def constraint_ode_ts1(self, ode, /, taylor_point=None):
        """ssm_impl_api.py:536-558 (prior-mean Taylor point).

        Dense: the exact Jacobian (every handler materialises it).  Isotropic
        (ssm_impl_isotropic.py:321-355): the trace over d of the Jacobian; blockdiag
        (ssm_impl_blockdiag.py:147-183): its diagonal.  Both are exact with jacobian_materialize
        (jacobians.py:100-113) and Hutchinson estimates with jacobian_monte_carlo_rev (ode()'s
        default) or jacobian_monte_carlo_fwd."""
        if not isinstance(ode, JetOde):
            raise TypeError(ode)
        if taylor_point is not None and not isinstance(taylor_point, taylor_point_prior):
            raise NotImplementedError("only the prior-mean Taylor point (taylor_point_prior) is implemented")
        _jacobian_handler(ode.jacobian)  # unknown handlers and probe counts fail here
        return Constraint("ts1", ode)


class state_space_model_dense(_StateSpaceModel):
    factorisation = "dense"


class state_space_model_isotropic(_StateSpaceModel):
    factorisation = "isotropic"


class state_space_model_blockdiag(_StateSpaceModel):
    factorisation = "blockdiag"


# -------------------------------------------------------- strategies/solvers
class _Strategy:
    name = ""
    is_suitable_for_save_at = True
    is_suitable_for_save_every_step = True
    is_suitable_for_offgrid_marginals = True

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}()"


class strategy_filter(_Strategy):
    """estimators_and_losses.py:347-421."""

    name = "filter"


class strategy_smoother_fixedpoint(_Strategy):
    """estimators_and_losses.py:473-591."""

    name = "fixedpoint"
    is_suitable_for_save_every_step = False
    is_suitable_for_offgrid_marginals = False


class strategy_smoother_fixedinterval(_Strategy):
    """estimators_and_losses.py:594-717."""

    name = "fixedinterval"
    is_suitable_for_save_at = False


class _Solver:
    kind = ""

    def __init__(self, *, strategy, constraint, constraint_init=None) -> None:
        if constraint_init is not None:
            raise NotImplementedError("constraint_init is not implemented on the device path")
        self.strategy = strategy
        self.constraint = constraint
        self.constraint_init = constraint_init

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}(strategy={self.strategy!r}, constraint={self.constraint!r})"

    @property
    def is_suitable_for_save_at(self):
        return self.strategy.is_suitable_for_save_at

    @property
    def is_suitable_for_save_every_step(self):
        return self.strategy.is_suitable_for_save_every_step

    @property
    def is_suitable_for_offgrid_marginals(self):
        return self.strategy.is_suitable_for_offgrid_marginals

    def offgrid_marginals(self, t, *, solution):
        """solvers.py:149-203: the marginal at t, strictly inside the solution's time interval and off
        its grid points (the reference does not check this either).  `t` may be a scalar or hold
        several times (a trailing axis; with a batched solution, a leading instance axis too): the
        device evaluates them all in one launch (the reference vmaps the scalar call)."""
        if not self.is_suitable_for_offgrid_marginals:
            raise NotImplementedError
        cfg = solution.config
        if cfg is None:
            raise ValueError("the solution does not carry the configuration of the solve that made it")
        batched = solution.t.dim() == 2
        t_all = solution.t if batched else solution.t[None]
        B, K = t_all.shape
        tq = _as_f64(t, t_all.device)
        scalar = tq.dim() == 0
        if tq.dim() <= 1:
            tq = tq.reshape(1, -1).expand(B, -1)
        tq = tq.contiguous()

        def flat(x, width):
            x = x if batched else x[None]
            return x.reshape(B, K, width).contiguous()

        N, d, ms, cs = cfg.N, cfg.ode_dim, cfg.mean_size, cfg.chol_size
        out = {"mean": flat(solution.u.mean_flat, ms), "chol": flat(solution.u.cholesky_flat, cs)}
        os_ = solution.output_scale if batched else solution.output_scale[None]
        os_ = os_.reshape(B, -1)
        if os_.shape[1] == K - 1:  # non-dynamic solvers: K - 1 rows; the reference's index K - 1 clamps
            os_ = torch.cat([os_, os_[:, -1:]], dim=1)
        out["output_scale"] = os_.contiguous()
        if self.strategy.name == "fixedinterval":
            filt = solution.solution_full.filtering
            out["filt_mean"] = flat(filt.mean_flat, ms)
            out["filt_chol"] = flat(filt.cholesky_flat, cs)
        mean, chol = capi.offgrid_marginals(cfg, out, t_all.contiguous(), tq)
        lead = (B, tq.shape[1]) if batched else (tq.shape[1],)
        mshape, lshape = ((cfg.n,), (cfg.n, cfg.n)) if cfg.factorisation == "dense" else ((N, d), (N, N))
        m = (mean if batched else mean[0]).reshape(*lead, *mshape)
        L = (chol if batched else chol[0]).reshape(*lead, *lshape)
        if scalar:
            m, L = m[..., 0, :] if cfg.factorisation == "dense" else m[..., 0, :, :], L[..., 0, :, :]
        return Normal(m, L, cfg.factorisation, N, d)


class solver(_Solver):
    """solvers.py:636-767 (no calibration)."""

    kind = "solver"


class solver_dynamic(_Solver):
    """solvers.py:483-633 (dynamic output-scale calibration)."""

    kind = "dynamic"

    def __init__(self, *, strategy, constraint, constraint_init=None, re_linearize_after_calibration=False, stop_gradient_through_calibration=True):
        super().__init__(strategy=strategy, constraint=constraint, constraint_init=constraint_init)
        self.re_linearize_after_calibration = bool(re_linearize_after_calibration)
        self.stop_gradient_through_calibration = stop_gradient_through_calibration


class solver_mle(_Solver):
    """solvers.py:318-480 (running maximum-likelihood calibration)."""

    kind = "mle"

    def __init__(self, *, strategy, constraint, constraint_init=None, correct_asymptotic_underconfidence=True):
        super().__init__(strategy=strategy, constraint=constraint, constraint_init=constraint_init)
        self.correct_asymptotic_underconfidence = correct_asymptotic_underconfidence


def error_norm_scale_then_rms(*, norm_order=None):
    """solvers.py:770-791."""
    if norm_order is not None:
        raise NotImplementedError("only the 2-norm is implemented")
    return "scale_then_rms"


def error_norm_rms_then_scale(norm_order=None):
    """solvers.py:794-811."""
    if norm_order is not None:
        raise NotImplementedError("only the 2-norm is implemented")
    return "rms_then_scale"


class error_residual_std:
    """solvers.py:850-996."""

    kind = "residual_std"
    derivative_idx = 0

    def __init__(self, *, constraint, error_norm=None, re_linearize_before_error=False, error_per_unit_step=False) -> None:
        self.constraint = constraint
        self.error_norm = error_norm or "scale_then_rms"
        self.re_linearize_before_error = bool(re_linearize_before_error)
        self.error_per_unit_step = bool(error_per_unit_step)


class error_state_std:
    """solvers.py:999-1098 (marked highly experimental in the reference): the calibrated std of
    derivative ``derivative_idx`` of the extrapolation conditioned on the step's linearisation.
    Both device kernels implement it (dense and isotropic factorisations)."""

    kind = "state_std"

    def __init__(self, *, constraint, error_norm=None, re_linearize_before_error=False, derivative_idx=0, error_per_unit_step=False) -> None:
        self.constraint = constraint
        self.error_norm = error_norm or "scale_then_rms"
        self.re_linearize_before_error = bool(re_linearize_before_error)
        self.derivative_idx = int(derivative_idx)
        self.error_per_unit_step = bool(error_per_unit_step)


# ------------------------------------------------------------- device call
def make_config(prior: Prior, solver_obj: _Solver, error=None, control=None, clip_dt=False) -> capi.Config:
    """Pack the configuration objects into the C ABI's POD struct."""
    ode_ = solver_obj.constraint.ode
    if ode_.ode_dim != prior.ode_dim:
        raise ValueError(f"the vector field has dimension {ode_.ode_dim}, the prior {prior.ode_dim}")
    cfg = capi.Config(
        factorisation=prior.factorisation,
        num_derivatives=prior.num_derivatives,
        ode_dim=prior.ode_dim,
        constraint=solver_obj.constraint.kind,
        strategy=solver_obj.strategy.name,
        solver=solver_obj.kind,
        vector_field=ode_.name,
        n_params=ode_.params_batched().shape[-1],
        clip_dt=clip_dt,
        base_output_scale=prior.base_output_scale,
        init_std=prior.init_std,
        init_std_rows=prior.init_std_rows,
    )
    if prior.kind == "exponential":
        cfg.prior, cfg.prior_drift = "exponential", prior.drift
    if solver_obj.constraint.kind == "ts1":
        cfg.jacobian, cfg.jac_probes, cfg.jac_seed = _jacobian_handler(ode_.jacobian)
    if isinstance(solver_obj, solver_mle):
        cfg.correct_asymptotic_underconfidence = solver_obj.correct_asymptotic_underconfidence
    if isinstance(solver_obj, solver_dynamic):
        cfg.re_linearize_after_calibration = solver_obj.re_linearize_after_calibration
    if error is not None:
        cfg.error_norm = error.error_norm
        cfg.error_estimator = error.kind
        cfg.error_re_linearize = error.re_linearize_before_error
        cfg.error_derivative_idx = error.derivative_idx
        cfg.error_per_unit_step = error.error_per_unit_step
    if control is not None:
        cfg.control = control.kind
        cfg.safety, cfg.factor_min, cfg.factor_max = control.safety, control.factor_min, control.factor_max
        if control.kind == "proportional_integral":
            cfg.exponent_integral = control.exponent_integral
            cfg.exponent_proportional = control.exponent_proportional
    return cfg


def build_solution(cfg: capi.Config, out: dict, *, batched: bool) -> ProbabilisticSolution:
    """Wrap the C ABI's flat outputs into the reference's solution structure."""
    B, K = out["t"].shape
    N, d = cfg.N, cfg.ode_dim
    fact = cfg.factorisation

    def normal(m, L):
        lead = m.shape[:-1]
        if fact == "dense":
            return Normal(m, L.reshape(*lead, N * d, N * d), fact, N, d)
        if fact == "isotropic":
            return Normal(m.reshape(*lead, N, d), L.reshape(*lead, N, N), fact, N, d)
        return Normal(m.reshape(*lead, d, N), L.reshape(*lead, d, N, N), fact, N, d)

    u = normal(out["mean"], out["chol"])
    if cfg.strategy == "filter":
        full = u
    else:
        n = cfg.n
        cond = LatentConditional(
            out["cond_A"].reshape(B, K - 1, n, n) if fact != "blockdiag" else out["cond_A"].reshape(B, K - 1, d, N, N),
            normal(out["cond_mean"], out["cond_chol"]),
            out["cond_tl"],
            out["cond_to"],
        )
        post = MarkovSequence(normal(out["term_mean"], out["term_chol"]), cond, reverse=True)
        full = SmoothingSolution(post, normal(out["filt_mean"], out["filt_chol"]))
    os_ = out["output_scale"]
    if cfg.solver != "dynamic":
        os_ = os_[:, : K - 1]
    if fact != "blockdiag":
        os_ = os_[..., 0]
    sol = ProbabilisticSolution(out["t"], u, full, os_, out["num_steps"], out["attempts"], out["status"], cfg)
    return sol if batched else sol.squeeze()


# ---------------------------------------------------------------- the LML
def loss_lml_terminal_values(*, tcoeff_index=0):
    """estimators_and_losses.py:20-50 on the device: the log-density of the data u under the
    marginals' derivative `tcoeff_index` observed with noise `std`.  ``marginals`` is an isotropic or
    dense Normal (e.g. a terminal-value solution's ``u``), single or batched; ``u`` has shape (d,) or
    (B, d); ``std`` is shaped as ``marginals.std[tcoeff_index]`` (a scalar isotropic, d dense)."""

    def loss(u, /, *, marginals, std):
        if marginals.factorisation not in ("isotropic", "dense"):
            raise NotImplementedError("loss_lml_terminal_values is implemented for the isotropic and dense factorisations")
        dev = marginals.mean_flat.device
        u_t = _as_f64(u, dev)
        std_t = _as_f64(std, dev)
        std_expected = marginals.std[tcoeff_index]
        if tuple(std_t.shape) != tuple(std_expected.shape):
            msg = "The standard deviation container differs from what was expected."
            msg += f" Expected: shape={tuple(std_expected.shape)}."
            msg += f" Received: shape={tuple(std_t.shape)}."
            msg += f" For reference: data-shape={tuple(u_t.shape)}."
            raise ValueError(msg)
        N, d = marginals.num_coeffs, marginals.ode_dim
        dense = marginals.factorisation == "dense"
        batched = marginals.mean_flat.dim() == (2 if dense else 3)
        m = marginals.mean_flat if batched else marginals.mean_flat[None]
        L = marginals.cholesky_flat if batched else marginals.cholesky_flat[None]
        B = m.shape[0]
        cfg = capi.Config(factorisation=marginals.factorisation, num_derivatives=N - 1, ode_dim=d)
        std_c = std_t.reshape(B, d) if dense else std_t.reshape(B)
        value = capi.lml_terminal_values(cfg, m.contiguous(), L.contiguous(), u_t.reshape(B, d).contiguous(), std_c.contiguous(), tcoeff_index)
        return value if batched else value[0]

    return loss


class linalg:  # noqa: N801 (the reference's module name)
    """The two solves loss_lml_timeseries accepts as ``solve_triu`` (backend/linalg.py:48-61):
    the device implements these and no other callable."""

    lstsq_svd = "lstsq_svd"
    solve_triu = "solve_triu"


def loss_lml_timeseries(*, average_pdfs: bool = True, tcoeff_index=0, solve_triu=linalg.lstsq_svd):
    """estimators_and_losses.py:53-105 on the device, as a differentiable torch op.

    Returns ``loss(u, /, *, problem, std)`` where ``problem`` is an
    ``LMLProblem`` (prior configuration, solver, grid and checkpoint stride):
    the reference differentiates through a solve that depends on the
    parameters; here the solve and the likelihood run in one launch and the
    gradient with respect to the vector-field parameters and the initial
    value comes from the same launch (forward-mode tangents, or the reverse
    adjoint with ``problem.lml_gradient = "reverse"``).  ``solve_triu`` is
    ``linalg.lstsq_svd`` (the reference's default) or ``linalg.solve_triu``;
    any other callable raises NotImplementedError.
    """
    tcoeff_index = int(tcoeff_index)
    if solve_triu is None:
        solve_triu = linalg.lstsq_svd
    if solve_triu not in (linalg.lstsq_svd, linalg.solve_triu):
        raise NotImplementedError("solve_triu must be linalg.lstsq_svd or linalg.solve_triu (the device's two gain solves)")

    def loss(u, /, *, problem, params, y0, std):
        from probdiffeq_amd.torch_ops import lml

        if not 0 <= tcoeff_index <= problem.num_derivatives:
            raise ValueError(f"tcoeff_index must lie in [0, {problem.num_derivatives}]")
        return lml(problem, u, std, params=params, y0=y0, average_pdfs=average_pdfs, tcoeff_index=tcoeff_index, solve=solve_triu)

    return loss

