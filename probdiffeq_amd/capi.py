"""ctypes binding of libpdq.so (the C ABI in include/pdq.h).

Device memory and streams come from PyTorch-ROCm: every pointer passed here
is a ``torch.Tensor.data_ptr()`` of a CUDA(HIP) tensor.  There is no CPU
path: without the built library or without a HIP device every call raises.
"""

from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Any

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libpdq.so")

FACT = {"dense": 0, "isotropic": 1, "blockdiag": 2}
CONSTRAINT = {"ts0": 0, "ts1": 1}
STRATEGY = {"filter": 0, "fixedpoint": 1, "fixedinterval": 2}
SOLVER = {"solver": 0, "dynamic": 1, "mle": 2}
ERROR_NORM = {"scale_then_rms": 0, "rms_then_scale": 1}
CONTROL = {"integral": 0, "proportional_integral": 1}
ERROR_ESTIMATOR = {"residual_std": 0, "state_std": 1}
PRIOR = {"iwp": 0, "exponential": 1}
LML_GRADIENT = {"forward": 0, "reverse": 1}
LML_SOLVE = {"lstsq_svd": 0, "solve_triu": 1}
JACOBIAN = {"exact": 0, "monte_carlo_rev": 1, "monte_carlo_fwd": 2}
VF = {"lotka_volterra": 0, "van_der_pol": 1, "burgers": 2, "linear": 3, "logistic": 4}
STATUS = {0: "ok", 1: "max_attempts", 2: "nonfinite", 3: "dt_underflow"}

ERR_NO_DEVICE = 3


class PdqError(RuntimeError):
    """A C-ABI call failed (the message comes from pdq_last_error)."""


class _Config(ctypes.Structure):
    _fields_ = [
        ("factorisation", ctypes.c_int32),
        ("num_derivatives", ctypes.c_int32),
        ("ode_dim", ctypes.c_int32),
        ("constraint", ctypes.c_int32),
        ("strategy", ctypes.c_int32),
        ("solver", ctypes.c_int32),
        ("error_norm", ctypes.c_int32),
        ("control", ctypes.c_int32),
        ("vector_field", ctypes.c_int32),
        ("n_params", ctypes.c_int32),
        ("clip_dt", ctypes.c_int32),
        ("max_attempts", ctypes.c_int32),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("dt0", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("damp", ctypes.c_double),
        ("safety", ctypes.c_double),
        ("factor_min", ctypes.c_double),
        ("factor_max", ctypes.c_double),
        ("exponent_integral", ctypes.c_double),
        ("exponent_proportional", ctypes.c_double),
        ("base_output_scale", ctypes.c_double),
        ("init_std", ctypes.c_double),
        ("correct_asymptotic_underconfidence", ctypes.c_int32),
        ("lanes_per_instance", ctypes.c_int32),
        ("re_linearize_after_calibration", ctypes.c_int32),
        ("error_estimator", ctypes.c_int32),
        ("error_re_linearize", ctypes.c_int32),
        ("error_derivative_idx", ctypes.c_int32),
        ("error_per_unit_step", ctypes.c_int32),
        ("prior", ctypes.c_int32),
        ("prior_drift", ctypes.c_void_p),
        ("dt0_per_instance", ctypes.c_void_p),
        ("lml_tcoeff_index", ctypes.c_int32),
        ("init_std_rows", ctypes.c_void_p),
        ("lml_gradient", ctypes.c_int32),
        ("lml_solve", ctypes.c_int32),
        ("jacobian", ctypes.c_int32),
        ("jac_probes", ctypes.c_int32),
        ("jac_seed", ctypes.c_uint32),
    ]


_VP = ctypes.c_void_p


class _Solution(ctypes.Structure):
    _fields_ = [
        (name, _VP)
        for name in [
            "t",
            "mean",
            "chol",
            "output_scale",
            "num_steps",
            "filt_mean",
            "filt_chol",
            "cond_A",
            "cond_mean",
            "cond_chol",
            "cond_tl",
            "cond_to",
            "term_mean",
            "term_chol",
            "attempts",
            "accepted",
            "status",
        ]
    ]


@dataclasses.dataclass
class Config:
    """Flat configuration mirroring the reference's solver objects."""

    factorisation: str = "dense"
    num_derivatives: int = 4
    ode_dim: int = 2
    constraint: str = "ts0"
    strategy: str = "filter"
    solver: str = "dynamic"
    error_norm: str = "scale_then_rms"
    control: str = "integral"
    vector_field: str = "lotka_volterra"
    n_params: int = 4
    clip_dt: bool = False
    max_attempts: int = 0
    rtol: float = 1e-6
    atol: float = 1e-8
    dt0: float = 0.1
    eps: float = 1e-8
    damp: float = 0.0
    safety: float = 0.95
    factor_min: float = 0.2
    factor_max: float = 10.0
    exponent_integral: float = 0.3
    exponent_proportional: float = 0.4
    base_output_scale: float = 1.0
    init_std: float = 0.0
    correct_asymptotic_underconfidence: bool = True
    lanes_per_instance: int = 0
    re_linearize_after_calibration: bool = False
    error_estimator: str = "residual_std"
    error_re_linearize: bool = False
    error_derivative_idx: int = 0
    error_per_unit_step: bool = False
    # "iwp", or "exponential" (dense): prior_drift holds the drift's last d rows, d x (nu+1) d
    # float64 (a device tensor for the library; host memory for the CPU build of the kernels)
    prior: str = "iwp"
    prior_drift: Any = None
    # optional initial step per instance: an (n_inst,) float64 device tensor (host array for the
    # CPU build of the kernels); None: dt0 for every instance
    dt0_per_instance: Any = None
    # loss_lml_timeseries(tcoeff_index=): the Taylor coefficient the LML's data observe
    lml_tcoeff_index: int = 0
    # optional initial std per covariance row ((nu+1) d dense, nu+1 isotropic): a float64 device
    # tensor (host array for the CPU build of the kernels); None: init_std on every row
    init_std_rows: Any = None
    # pdq_lml_value_and_grad's gradient: "forward" (tangents) or "reverse" (the adjoint)
    lml_gradient: str = "forward"
    # loss_lml_timeseries(solve_triu=): "lstsq_svd" (the reference's default) or "solve_triu"
    lml_solve: str = "lstsq_svd"
    # ts1 Jacobian handler: "exact" (jacobian_materialize) or "monte_carlo_rev"
    # or "monte_carlo_fwd" (jacobian_monte_carlo_rev / _fwd(seed=jac_seed, num_probes=jac_probes));
    # isotropic and blockdiag only (the dense factorisation takes the exact Jacobian)
    jacobian: str = "exact"
    jac_probes: int = 10
    jac_seed: int = 1

    def c(self, host: bool = False) -> _Config:
        """The C struct.  The kernels dereference init_std_rows, prior_drift and
        dt0_per_instance on the device, so for the HIP library (host=False) they
        must be contiguous float64 device tensors; host=True (the CPU build of the
        kernels) takes numpy arrays instead."""
        c = self._c()
        c.lml_tcoeff_index = int(self.lml_tcoeff_index)
        if self.lml_gradient not in LML_GRADIENT:
            raise ValueError(f"unknown lml_gradient {self.lml_gradient!r}")
        c.lml_gradient = LML_GRADIENT[self.lml_gradient]
        if self.lml_solve not in LML_SOLVE:
            raise ValueError(f"unknown lml_solve {self.lml_solve!r}")
        c.lml_solve = LML_SOLVE[self.lml_solve]
        if self.jacobian not in JACOBIAN:
            raise ValueError(f"unknown jacobian {self.jacobian!r}")
        c.jacobian = JACOBIAN[self.jacobian]
        c.jac_probes = int(self.jac_probes)
        c.jac_seed = int(self.jac_seed) & 0xFFFFFFFF
        n_rows = (self.num_derivatives + 1) * (self.ode_dim if self.factorisation == "dense" else 1)
        if self.prior not in PRIOR:
            raise ValueError(f"unknown prior {self.prior!r}")
        c.prior = PRIOR[self.prior]
        for field, shape in (
            ("init_std_rows", (n_rows,)),
            ("prior_drift", (self.ode_dim, (self.num_derivatives + 1) * self.ode_dim)),
            ("dt0_per_instance", None),
        ):
            v = getattr(self, field)
            if v is None:
                continue
            if shape is not None and tuple(v.shape) != shape:
                raise ValueError(f"{field} must have shape {shape} (got {tuple(v.shape)})")
            if v.ndim != (1 if shape is None else len(shape)):
                raise ValueError(f"{field} has {v.ndim} dimensions")
            if host:
                if not isinstance(v, np.ndarray):
                    raise ValueError(f"{field}: the CPU build of the kernels takes a numpy array")
                v = np.ascontiguousarray(v, dtype=np.float64)
                ptr = v.ctypes.data
            else:
                if not isinstance(v, torch.Tensor) or not v.is_cuda:
                    raise PdqError(f"{field} must be a device (HIP) tensor: the kernels read it on the GPU")
                if v.dtype != torch.float64 or not v.is_contiguous():
                    raise ValueError(f"{field} must be a contiguous float64 tensor")
                ptr = v.data_ptr()
            setattr(c, field, ptr)
            setattr(c, "_keep_" + field, v)  # the memory lives as long as the struct
        return c

    def _c(self) -> _Config:
        return _Config(
            FACT[self.factorisation],
            self.num_derivatives,
            self.ode_dim,
            CONSTRAINT[self.constraint],
            STRATEGY[self.strategy],
            SOLVER[self.solver],
            ERROR_NORM[self.error_norm],
            CONTROL[self.control],
            VF[self.vector_field],
            self.n_params,
            int(self.clip_dt),
            self.max_attempts,
            self.rtol,
            self.atol,
            self.dt0,
            self.eps,
            self.damp,
            self.safety,
            self.factor_min,
            self.factor_max,
            self.exponent_integral,
            self.exponent_proportional,
            self.base_output_scale,
            self.init_std,
            int(self.correct_asymptotic_underconfidence),
            self.lanes_per_instance,
            int(self.re_linearize_after_calibration),
            ERROR_ESTIMATOR[self.error_estimator],
            int(self.error_re_linearize),
            self.error_derivative_idx,
            int(self.error_per_unit_step),
        )

    @property
    def N(self) -> int:
        return self.num_derivatives + 1

    @property
    def nblk(self) -> int:
        return self.ode_dim if self.factorisation == "blockdiag" else 1

    @property
    def n(self) -> int:
        return self.N * self.ode_dim if self.factorisation == "dense" else self.N

    @property
    def mean_size(self) -> int:
        return self.N * self.ode_dim

    @property
    def chol_size(self) -> int:
        return self.nblk * self.n * self.n


# Symbols the header declares; the CPU test suite checks that all exist.
EXPORTED = [
    "pdq_version",
    "pdq_system_matrices_1d_iwp",
    "pdq_init",
    "pdq_last_error",
    "pdq_solve_adaptive_save_at",
    "pdq_solve_adaptive_save_at_noalloc",
    "pdq_solve_fixed_grid",
    "pdq_lml_value_and_grad",
    "pdq_taylor_coefficients",
    "pdq_mean_size",
    "pdq_chol_size",
    "pdq_tuned_lanes",
    "pdq_set_tuned_lanes",
    "pdq_offgrid_marginals",
    "pdq_lml_terminal_values",
    "pdq_sample",
]

_lib = None


def lib() -> ctypes.CDLL:
    """Load libpdq.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PdqError(f"{LIB_PATH} is missing: build it with probdiffeq_amd.build.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name in EXPORTED:
            getattr(L, name).restype = ctypes.c_int32
        L.pdq_last_error.restype = ctypes.c_size_t
        _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().pdq_last_error(buf, 1024)
    return buf.value.decode()


def check(rc: int) -> None:
    if rc != 0:
        raise PdqError(f"pdq error {rc}: {last_error()}")


def _ptr(t, dtype=torch.float64, shape=None, name="tensor"):
    """Device pointer of a contiguous tensor of the given dtype (and shape,
    where given: None entries match any size)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise PdqError(f"pdq takes device tensors (HIP); {name} is not one")
    if not t.is_contiguous():
        raise PdqError(f"pdq takes contiguous tensors ({name} is not)")
    if dtype is not None and t.dtype != dtype:
        raise PdqError(f"{name} must be {dtype} (got {t.dtype})")
    if shape is not None:
        if t.dim() != len(shape) or any(s is not None and s != n for s, n in zip(shape, t.shape)):
            raise PdqError(f"{name} must have shape {tuple('*' if s is None else s for s in shape)} (got {tuple(t.shape)})")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def require_device() -> torch.device:
    if not torch.cuda.is_available():
        raise PdqError("no HIP device: the solver path runs on MI355X (gfx950) only; there is no CPU fallback")
    return torch.device("cuda")


def alloc_solution(cfg: Config, n_inst: int, K: int, device=None, *, smoother_storage: bool | None = None):
    """Allocate the output tensors of one batched solve (see pdq_solution)."""
    dev = device or require_device()
    f64 = dict(dtype=torch.float64, device=dev)
    ms, cs, nn, nv = cfg.mean_size, cfg.chol_size, cfg.nblk * cfg.n * cfg.n, cfg.nblk * cfg.n
    out = {
        "t": torch.zeros(n_inst, K, **f64),
        "mean": torch.zeros(n_inst, K, ms, **f64),
        "chol": torch.zeros(n_inst, K, cs, **f64),
        "output_scale": torch.zeros(n_inst, K, cfg.nblk, **f64),
        "num_steps": torch.zeros(n_inst, K - 1, dtype=torch.int64, device=dev),
        "attempts": torch.zeros(n_inst, dtype=torch.int64, device=dev),
        "accepted": torch.zeros(n_inst, dtype=torch.int64, device=dev),
        "status": torch.zeros(n_inst, dtype=torch.int32, device=dev),
    }
    smoother = cfg.strategy != "filter" if smoother_storage is None else smoother_storage
    if smoother:
        out.update(
            filt_mean=torch.zeros(n_inst, K, ms, **f64),
            filt_chol=torch.zeros(n_inst, K, cs, **f64),
            cond_A=torch.zeros(n_inst, K - 1, nn, **f64),
            cond_mean=torch.zeros(n_inst, K - 1, ms, **f64),
            cond_chol=torch.zeros(n_inst, K - 1, cs, **f64),
            cond_tl=torch.zeros(n_inst, K - 1, nv, **f64),
            cond_to=torch.zeros(n_inst, K - 1, nv, **f64),
            term_mean=torch.zeros(n_inst, ms, **f64),
            term_chol=torch.zeros(n_inst, cs, **f64),
        )
    return out


_INT_OUTPUTS = {"num_steps": torch.int64, "attempts": torch.int64, "accepted": torch.int64, "status": torch.int32}


def _solution_struct(out: dict) -> _Solution:
    s = _Solution()
    for name, _ in _Solution._fields_:
        t = out.get(name)
        setattr(s, name, _ptr(t, _INT_OUTPUTS.get(name, torch.float64), name=name).value if t is not None else None)
    return s


def solve(cfg: Config, params, tcoeffs, times, *, fixed_grid=False, out=None, stream=None, no_alloc=False):
    """Batched solve; params (B, P), tcoeffs (B, N, d), times (B, K) device tensors.

    no_alloc=True calls pdq_solve_adaptive_save_at_noalloc (save_at only): the
    library allocates no device memory for the call (out, if given, is the
    caller's)."""
    if times.dim() != 2:
        raise PdqError(f"times must have shape (B, K) (got {tuple(times.shape)})")
    n_inst, K = times.shape
    p_ptr = _ptr(params, shape=(n_inst, cfg.n_params), name="params")
    tc_ptr = _ptr(tcoeffs, shape=(n_inst, cfg.N, cfg.ode_dim), name="tcoeffs")
    t_ptr = _ptr(times, name="times")
    if out is None:
        out = alloc_solution(cfg, n_inst, K, times.device)
    c = cfg.c()
    s = _solution_struct(out)
    if no_alloc:
        if fixed_grid:
            raise PdqError("the no-allocation entry point solves on save_at points (pdq_solve_adaptive_save_at_noalloc)")
        fn = lib().pdq_solve_adaptive_save_at_noalloc
    else:
        fn = lib().pdq_solve_fixed_grid if fixed_grid else lib().pdq_solve_adaptive_save_at
    check(fn(ctypes.byref(c), n_inst, p_ptr, tc_ptr, t_ptr, K, ctypes.byref(s), _stream(stream)))
    return out


def lml_value_and_grad(cfg: Config, params, y0, grid, every, data, stds, *, average=True, grad=True, stream=None):
    """Batched marginal log-likelihood (and its gradient w.r.t. params and y0)."""
    n_inst = params.shape[0]
    dev = params.device
    if grid.dim() != 1 or every < 1 or (grid.shape[0] - 1) % every != 0:
        raise PdqError(f"grid must be (K_grid,) with (K_grid - 1) divisible by every = {every}")
    k_obs = (grid.shape[0] - 1) // every + 1
    n_std = stds.shape[-1] if stds.dim() == 3 else 1
    args = (
        _ptr(params, shape=(n_inst, cfg.n_params), name="params"),
        _ptr(y0, shape=(n_inst, cfg.ode_dim), name="y0"),
        _ptr(grid, name="grid"),
        _ptr(data, shape=(n_inst, k_obs, cfg.ode_dim), name="data"),
        _ptr(stds, shape=(n_inst, k_obs, n_std) if stds.dim() == 3 else (n_inst, k_obs), name="stds"),
    )
    value = torch.zeros(n_inst, dtype=torch.float64, device=dev)
    g = torch.zeros(n_inst, cfg.n_params + cfg.ode_dim, dtype=torch.float64, device=dev) if grad else None
    c = cfg.c()
    p, y, gr, da, st = args
    check(
        lib().pdq_lml_value_and_grad(
            ctypes.byref(c), n_inst, p, y, gr, grid.shape[0], every, da, st, n_std, int(average), _ptr(value), _ptr(g), _stream(stream)
        )
    )
    return value, g


def offgrid_marginals(cfg: Config, out: dict, times, t_query, stream=None):
    """Marginals at off-grid times t_query (B, Q) of the solution `out` that a solve with `cfg`
    wrote on `times` (B, K) (solvers.py:149-203).  Returns (mean (B, Q, mean_size),
    chol (B, Q, chol_size))."""
    B, K = times.shape
    Qn = t_query.shape[1]
    mean = torch.empty(B, Qn, cfg.mean_size, dtype=torch.float64, device=times.device)
    chol = torch.empty(B, Qn, cfg.chol_size, dtype=torch.float64, device=times.device)
    c = cfg.c()
    s = _solution_struct(out)
    check(lib().pdq_offgrid_marginals(ctypes.byref(c), B, K, _ptr(times), ctypes.byref(s), Qn, _ptr(t_query), _ptr(mean), _ptr(chol), _stream(stream)))
    return mean, chol


def sample(cfg: Config, out: dict, K: int, base, stream=None):
    """Posterior samples (estimators_and_losses.py:233-273) of the smoother solution `out` that a solve
    with `cfg` wrote on K points, from the standard normal draws base (B, S, K, N).  Returns
    samples (B, S, K, mean_size)."""
    B, S = base.shape[0], base.shape[1]
    samples = torch.empty(B, S, K, cfg.mean_size, dtype=torch.float64, device=base.device)
    c = cfg.c()
    s = _solution_struct(out)
    check(lib().pdq_sample(ctypes.byref(c), B, K, ctypes.byref(s), S, _ptr(base), _ptr(samples), _stream(stream)))
    return samples


def lml_terminal_values(cfg: Config, mean, chol, u, std, tcoeff_index=0, stream=None):
    """loss_lml_terminal_values (estimators_and_losses.py:20-50) for a batch of marginals: isotropic
    mean (B, N, d), chol (B, N, N), std (B,); dense mean (B, N d), chol (B, N d, N d), std (B, d);
    data u (B, d).  Returns (B,) log-densities."""
    B = mean.shape[0]
    value = torch.empty(B, dtype=torch.float64, device=mean.device)
    c = cfg.c()
    check(lib().pdq_lml_terminal_values(ctypes.byref(c), B, _ptr(mean), _ptr(chol), _ptr(u), _ptr(std), int(tcoeff_index), _ptr(value), _stream(stream)))
    return value


def system_matrices_1d_iwp(num_derivatives: int):
    """The library's latent IWP matrices (A, Q), (nu+1) x (nu+1) numpy arrays (host only)."""
    n = num_derivatives + 1
    A, Q = np.zeros((n, n)), np.zeros((n, n))
    check(lib().pdq_system_matrices_1d_iwp(num_derivatives, A.ctypes.data_as(ctypes.c_void_p), Q.ctypes.data_as(ctypes.c_void_p)))
    return A, Q


def taylor_coefficients(vector_field: str, params, y0, t0: float, num: int, stream=None):
    """Initial Taylor coefficients (derivatives) of the registered vector field."""
    n_inst, d = y0.shape
    out = torch.zeros(n_inst, num + 1, d, dtype=torch.float64, device=y0.device)
    check(
        lib().pdq_taylor_coefficients(
            VF[vector_field],
            d,
            params.shape[1],
            n_inst,
            _ptr(params, shape=(n_inst, None), name="params"),
            _ptr(y0, name="y0"),
            ctypes.c_double(t0),
            num,
            _ptr(out),
            _stream(stream),
        )
    )
    return out


def set_tuned_lanes(cfg: Config, n_inst: int, lanes: int, mode: int = 0) -> None:
    c = cfg.c()
    check(lib().pdq_set_tuned_lanes(ctypes.byref(c), n_inst, mode, lanes))
