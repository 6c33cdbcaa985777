"""ctypes front-end of the CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

The oracle restates the reference's JAX solver path in C++
(``oracle/pdq_oracle.cpp``).  Only ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` may import this module; the product
package never does.
"""

from __future__ import annotations

import contextlib
import ctypes
import dataclasses
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpdq_oracle.so")
_LIB_PATHS = {
    "strict": _LIB_PATH,
    "fma": os.path.join(_HERE, "_build", "libpdq_oracle_fma.so"),
    "extended": os.path.join(_HERE, "_build", "libpdq_oracle_ext.so"),
    # the strict build at -O3 for x86-64-v3 (AVX2, FMA units; contraction still off, so the
    # arithmetic and the step sequence are the strict build's): the CPU baseline's timing leg
    "fast": os.path.join(_HERE, "_build", "libpdq_oracle_fast.so"),
}
# scalar type of each build: (numpy dtype, ctypes type)
_REAL = {
    "strict": (np.float64, ctypes.c_double),
    "fma": (np.float64, ctypes.c_double),
    "extended": (np.longdouble, ctypes.c_longdouble),
    "fast": (np.float64, ctypes.c_double),
}

FACT = {"dense": 0, "isotropic": 1, "blockdiag": 2}
CONSTRAINT = {"ts0": 0, "ts1": 1}
STRATEGY = {"filter": 0, "fixedpoint": 1, "fixedinterval": 2}
SOLVER = {"solver": 0, "dynamic": 1, "mle": 2}
ERROR_NORM = {"scale_then_rms": 0, "rms_then_scale": 1}
CONTROL = {"integral": 0, "proportional_integral": 1}
JACOBIAN = {"exact": 0, "monte_carlo_rev": 1, "monte_carlo_fwd": 2}
VF = {"lotka_volterra": 0, "van_der_pol": 1, "burgers": 2, "linear": 3, "logistic": 4}


def _cfg_struct(ct):
    class _Cfg(ctypes.Structure):
        _fields_ = [
            ("fact", ctypes.c_int),
            ("nu", ctypes.c_int),
            ("d", ctypes.c_int),
            ("constraint", ctypes.c_int),
            ("strategy", ctypes.c_int),
            ("solver", ctypes.c_int),
            ("error_norm", ctypes.c_int),
            ("control", ctypes.c_int),
            ("vf_id", ctypes.c_int),
            ("clip_dt", ctypes.c_int),
            ("rtol", ct),
            ("atol", ct),
            ("dt0", ct),
            ("eps", ct),
            ("damp", ct),
            ("safety", ct),
            ("factor_min", ct),
            ("factor_max", ct),
            ("exponent_integral", ct),
            ("exponent_proportional", ct),
            ("base_scale", ct),
            ("correct_asymptotic_underconfidence", ctypes.c_int),
            ("max_attempts", ctypes.c_int),
            ("re_linearize", ctypes.c_int),
            ("error_estimator", ctypes.c_int),
            ("error_re_linearize", ctypes.c_int),
            ("error_derivative_idx", ctypes.c_int),
            ("error_per_unit_step", ctypes.c_int),
            ("prior", ctypes.c_int),
            ("prior_drift", ctypes.POINTER(ct)),
            ("tcoeff_index", ctypes.c_int),
            ("init_std_rows", ctypes.POINTER(ct)),
            ("lml_solve", ctypes.c_int),
            ("jacobian", ctypes.c_int),
            ("jac_probes", ctypes.c_int),
            ("jac_seed", ctypes.c_uint32),
        ]

    return _Cfg


def _out_struct(ct):
    P = ctypes.POINTER(ct)

    class _Out(ctypes.Structure):
        _fields_ = [
            ("t", P),
            ("mean", P),
            ("chol", P),
            ("os", P),
            ("num_steps", _PL),
            ("filt_mean", P),
            ("filt_chol", P),
            ("cond_A", P),
            ("cond_mean", P),
            ("cond_chol", P),
            ("cond_tl", P),
            ("cond_to", P),
            ("term_mean", P),
            ("term_chol", P),
            ("attempts", ctypes.c_long),
            ("accepted", ctypes.c_long),
            ("status", ctypes.c_int),
        ]

    return _Out


_PL = ctypes.POINTER(ctypes.c_long)
_CFG = {ct: _cfg_struct(ct) for ct in (ctypes.c_double, ctypes.c_longdouble)}
_OUT = {ct: _out_struct(ct) for ct in (ctypes.c_double, ctypes.c_longdouble)}


def _dtype():
    return _REAL[_variant][0]


def _ct():
    return _REAL[_variant][1]


def _zeros(shape):
    return np.zeros(shape, dtype=_dtype())


def _s(x):
    """A scalar argument in the active build's type."""
    return _ct()(x)


@dataclasses.dataclass
class Config:
    """Solver configuration (mirrors the reference's constructor kwargs)."""

    fact: str = "dense"
    nu: int = 4
    d: int = 2
    constraint: str = "ts0"
    strategy: str = "filter"
    solver: str = "dynamic"
    error_norm: str = "scale_then_rms"
    control: str = "integral"
    vf: str = "lotka_volterra"
    clip_dt: bool = False
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
    base_scale: float = 1.0
    correct_asymptotic_underconfidence: bool = True
    max_attempts: int = 0
    re_linearize_after_calibration: bool = False
    error_estimator: str = "residual_std"
    error_re_linearize: bool = False
    error_derivative_idx: int = 0
    error_per_unit_step: bool = False
    # "iwp" or "exponential" (dense): prior_drift is the drift's bottom block,
    # d x (nu+1) d, derivative-major columns (ssm_impl_dense.py:541-629)
    prior: str = "iwp"
    prior_drift: object = None
    # loss_lml_timeseries(tcoeff_index=): the Taylor coefficient the LML's data observe
    tcoeff_index: int = 0
    # optional initial std per covariance row (n values; dense (nu+1) d, isotropic nu+1)
    init_std_rows: object = None
    # loss_lml_timeseries(solve_triu=): "lstsq_svd" (the default) or "solve_triu"
    lml_solve: str = "lstsq_svd"
    # ts1 Jacobian handler: "exact" (jacobian_materialize), "monte_carlo_rev" or
    # "monte_carlo_fwd" (jacobian_monte_carlo_rev / _fwd(seed=jac_seed, num_probes=jac_probes))
    jacobian: str = "exact"
    jac_probes: int = 10
    jac_seed: int = 1

    def c(self):
        if self.prior not in ("iwp", "exponential"):
            raise ValueError(f"unknown prior {self.prior!r}")
        drift = None
        if self.prior == "exponential":
            if self.fact != "dense":
                raise ValueError("exponential priors are dense only (ssm_impl_isotropic.py:484-520)")
            drift = np.ascontiguousarray(np.asarray(self.prior_drift, dtype=np.float64).reshape(self.d, self.N * self.d), dtype=_dtype())
        c = self._c()
        c.prior = 1 if drift is not None else 0
        if not 0 <= self.tcoeff_index <= self.nu:
            raise ValueError("tcoeff_index must lie in [0, nu]")
        c.tcoeff_index = int(self.tcoeff_index)
        if self.lml_solve not in ("lstsq_svd", "solve_triu"):
            raise ValueError(f"unknown lml_solve {self.lml_solve!r}")
        c.lml_solve = 1 if self.lml_solve == "solve_triu" else 0
        if self.jacobian not in JACOBIAN:
            raise ValueError(f"unknown jacobian {self.jacobian!r}")
        if self.jacobian != "exact" and not 1 <= self.jac_probes <= 64:
            raise ValueError("jac_probes must lie in [1, 64]")
        c.jacobian, c.jac_probes, c.jac_seed = JACOBIAN[self.jacobian], int(self.jac_probes), int(self.jac_seed) & 0xFFFFFFFF
        if self.init_std_rows is not None:
            rows = np.ascontiguousarray(np.asarray(self.init_std_rows, dtype=np.float64), dtype=_dtype())
            if rows.shape != (self.n,):
                raise ValueError(f"init_std_rows must have shape ({self.n},)")
            c.init_std_rows = rows.ctypes.data_as(ctypes.POINTER(_ct()))
            c._keep_rows = rows
        if drift is not None:
            c.prior_drift = drift.ctypes.data_as(ctypes.POINTER(_ct()))
            c._keep = drift  # the array lives as long as the struct
        return c

    def _c(self):
        return _CFG[_ct()](
            FACT[self.fact],
            self.nu,
            self.d,
            CONSTRAINT[self.constraint],
            STRATEGY[self.strategy],
            SOLVER[self.solver],
            ERROR_NORM[self.error_norm],
            CONTROL[self.control],
            VF[self.vf],
            int(self.clip_dt),
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
            self.base_scale,
            int(self.correct_asymptotic_underconfidence),
            self.max_attempts,
            int(self.re_linearize_after_calibration),
            {"residual_std": 0, "state_std": 1}[self.error_estimator],
            int(self.error_re_linearize),
            int(self.error_derivative_idx),
            int(self.error_per_unit_step),
        )

    @property
    def N(self) -> int:
        return self.nu + 1

    @property
    def nblk(self) -> int:
        return self.d if self.fact == "blockdiag" else 1

    @property
    def n(self) -> int:
        return self.N * self.d if self.fact == "dense" else self.N

    @property
    def mean_size(self) -> int:
        return self.N * self.d

    @property
    def chol_size(self) -> int:
        return self.nblk * self.n * self.n


def build(force: bool = False) -> str:
    """Compile both oracle variants with the Makefile (g++, no GPU needed)."""
    if force or not all(os.path.exists(p) for p in _LIB_PATHS.values()):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


_libs = {}
_variant = "strict"


@contextlib.contextmanager
def rounding(name: str):
    """Run oracle calls inside the block on the given build ("strict" or "fma").

    The "fma" build is the same restatement compiled with FMA contraction; the
    difference between the two builds' outputs is the rounding sensitivity of
    each output, used to scale parity tolerances of ill-conditioned quantities.
    """
    global _variant
    if name not in _LIB_PATHS:
        raise ValueError(f"unknown oracle build {name!r}")
    prev, _variant = _variant, name
    try:
        yield
    finally:
        _variant = prev


def lib():
    L = _libs.get(_variant)
    if L is None:
        path = _LIB_PATHS[_variant]
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        ct = _REAL[_variant][1]
        if L.pdqo_real_size() != ctypes.sizeof(ct):
            raise RuntimeError(f"oracle build {_variant!r} has a {L.pdqo_real_size()}-byte scalar, expected {ctypes.sizeof(ct)}")
        L.pdqo_solve_save_at.restype = ctypes.c_int
        L.pdqo_solve_fixed_grid.restype = ctypes.c_int
        L.pdqo_lml.restype = ct
        # full-precision return: ctypes converts a plain c_longdouble result to a
        # Python float, but not an instance of a subclass (read back by _full)
        L.pdqo_lml_from_y0.restype = _LongDoubleResult if ct is ctypes.c_longdouble else ct
        L.pdqo_lml_from_y0_tape.restype = _LongDoubleResult if ct is ctypes.c_longdouble else ct
        L.pdqo_lml_grad_fd.restype = ct
        L.pdqo_lml_terminal.restype = ct
        L.pdqo_step_error_power.restype = ct
        L.pdqo_logpdf.restype = ct
        L.pdqo_control.restype = ct
        L.pdqo_rademacher.restype = ct
        L.pdqo_rademacher.argtypes = [ctypes.c_uint32] * 4
        L.pdqo_solve_save_at_batch.restype = ctypes.c_int
        _libs[_variant] = L
    return L


class _LongDoubleResult(ctypes.c_longdouble):
    """A long double return value that ctypes leaves unconverted."""


def _full(v):
    """A scalar oracle result at the build's precision (np.longdouble for the extended build)."""
    if isinstance(v, ctypes.c_longdouble):
        return np.frombuffer(bytes(v), dtype=np.longdouble)[0]
    return v


def _p(a):
    if a is None:
        return None
    assert a.dtype == _dtype(), (a.dtype, _dtype())
    return a.ctypes.data_as(ctypes.POINTER(_ct()))


def _f64(a):
    """Input array in the active build's scalar type (float64 or long double)."""
    return np.ascontiguousarray(np.asarray(a, dtype=_dtype()))


def qr_r(M):
    M = _f64(M)
    m, n = M.shape
    R = _zeros((min(m, n), n))
    lib().pdqo_qr_r(m, n, _p(M), _p(R))
    return R


def lstsq_svd(A, B):
    """jnp.linalg.lstsq(A, B)[0] (backend/linalg.py:60-61): the minimum-norm least-squares
    solution, singular values at or below eps * max(m, n) * s_max treated as zero."""
    A, B = _f64(A), _f64(B)
    vec = B.ndim == 1
    B = _f64(B.reshape(B.shape[0], -1))
    m, n = A.shape
    X = _zeros((n, B.shape[1]))
    lib().pdqo_lstsq_svd(m, n, _p(A), B.shape[1], _p(B), _p(X))
    return X[:, 0] if vec else X


def revert_conditional(R_X_F, R_X, R_YX, lstsq=False):
    R_X_F, R_X, R_YX = _f64(R_X_F), _f64(R_X), _f64(R_YX)
    n, k = R_X_F.shape
    R_Y, R_XY, G = _zeros((k, k)), _zeros((n, n)), _zeros((n, k))
    lib().pdqo_revert_conditional(n, k, _p(R_X_F), _p(R_X), _p(R_YX), int(lstsq), _p(R_Y), _p(R_XY), _p(G))
    return R_Y, (R_XY, G)


def system_matrices(nu):
    A = _zeros((nu + 1, nu + 1))
    Q = _zeros((nu + 1, nu + 1))
    lib().pdqo_system_matrices(nu, _p(A), _p(Q))
    return A, Q


def exp_gram(A, B):
    """gram_util.exp_gram_cholesky(pade_and_legendre_9): (e^A, U) with U U^T the Gramian."""
    A, B = _f64(A), _f64(B)
    n, m = B.shape
    eA, U = _zeros((n, n)), _zeros((n, n))
    lib().pdqo_exp_gram(n, m, _p(A), _p(B), _p(eA), _p(U))
    return eA, U


def transition(cfg: Config, dt, os=None):
    n = cfg.n
    os = _f64(np.ones(cfg.nblk) if os is None else np.atleast_1d(os))
    A, L, tl, to = _zeros((n, n)), _zeros((n, n)), _zeros(n), _zeros(n)
    c = cfg.c()
    lib().pdqo_transition(ctypes.byref(c), _s(dt), _p(os), _p(A), _p(L), _p(tl), _p(to))
    return A, L, tl, to


def logpdf(mean, chol, u, isotropic=False):
    mean, chol, u = _f64(mean), _f64(chol), _f64(u)
    if mean.ndim == 1:
        mean, u = mean[:, None], u[:, None]
    return lib().pdqo_logpdf(int(isotropic), mean.shape[0], mean.shape[1], _p(mean), _p(chol), _p(u))


def control(cfg: Config, dt, state, error_power):
    s = _s(state)
    c = cfg.c()
    dt_new = lib().pdqo_control(ctypes.byref(c), _s(dt), ctypes.byref(s), _s(error_power))
    return dt_new, s.value


def taylor_coefficients(vf, params, y0, t0, num):
    params, y0 = _f64(params), _f64(y0)
    d = y0.size
    out = _zeros((num + 1, d))
    lib().pdqo_taylor_coefficients(VF[vf], d, _p(params), _p(y0), _s(t0), num, _p(out))
    return out


def vector_field(vf, params, y, t=0.0):
    params, y = _f64(params), _f64(y)
    d = y.size
    f, J = _zeros(d), _zeros((d, d))
    lib().pdqo_vf(VF[vf], d, _p(params), _p(y), _s(t), _p(f), _p(J))
    return f, J


def rademacher(seed, set_, probe, coord):
    """One entry of the Monte Carlo handlers' probe stream (+1.0 or -1.0)."""
    return float(lib().pdqo_rademacher(seed & 0xFFFFFFFF, set_ & 0xFFFFFFFF, probe, coord))


def mc_jacobian_diagonal(vf, params, y, *, seed=1, set_=0, probes=10, mode="rev", t=0.0):
    """The Monte Carlo estimate of the vector field's Jacobian diagonal at y (calculate_diagonal_along_d)."""
    params, y = _f64(params), _f64(y)
    d = y.size
    out = _zeros(d)
    if mode not in ("rev", "fwd"):
        raise ValueError(f"unknown mode {mode!r}")
    lib().pdqo_mc_diag(VF[vf], d, _p(params), _p(y), _s(t), ctypes.c_uint32(seed & 0xFFFFFFFF),
                       ctypes.c_uint32(set_ & 0xFFFFFFFF), int(probes), int(mode == "fwd"), _p(out))
    return out


def _alloc(cfg: Config, K: int):
    ms, cs, nn, nv = cfg.mean_size, cfg.chol_size, cfg.nblk * cfg.n * cfg.n, cfg.nblk * cfg.n
    smoother = cfg.strategy != "filter"
    arrs = {
        "t": _zeros(K),
        "mean": _zeros((K, ms)),
        "chol": _zeros((K, cs)),
        "output_scale": _zeros((K, cfg.nblk)),
        "num_steps": np.zeros(K - 1, dtype=np.int64),
    }
    if smoother:
        arrs.update(
            filt_mean=_zeros((K, ms)),
            filt_chol=_zeros((K, cs)),
            cond_A=_zeros((K - 1, nn)),
            cond_mean=_zeros((K - 1, ms)),
            cond_chol=_zeros((K - 1, cs)),
            cond_tl=_zeros((K - 1, nv)),
            cond_to=_zeros((K - 1, nv)),
            term_mean=_zeros(ms),
            term_chol=_zeros(cs),
        )
    out = _OUT[_ct()]()
    out.t, out.mean, out.chol, out.os = _p(arrs["t"]), _p(arrs["mean"]), _p(arrs["chol"]), _p(arrs["output_scale"])
    out.num_steps = arrs["num_steps"].ctypes.data_as(_PL)
    if smoother:
        for k in ["filt_mean", "filt_chol", "cond_A", "cond_mean", "cond_chol", "cond_tl", "cond_to", "term_mean", "term_chol"]:
            setattr(out, k, _p(arrs[k]))
    return arrs, out


def _finish(cfg, arrs, out):
    arrs["attempts"] = out.attempts
    arrs["accepted"] = out.accepted
    arrs["status"] = out.status
    if cfg.solver != "dynamic":
        arrs["output_scale"] = arrs["output_scale"][:-1]
    return arrs


def solve_save_at(cfg: Config, params, tcoeffs, save_at, init_std=0.0):
    params, tcoeffs, save_at = _f64(params), _f64(tcoeffs), _f64(save_at)
    K = save_at.size
    arrs, out = _alloc(cfg, K)
    c = cfg.c()
    st = lib().pdqo_solve_save_at(ctypes.byref(c), _p(params), _p(tcoeffs), _s(init_std), _p(save_at), K, ctypes.byref(out))
    arrs = _finish(cfg, arrs, out)
    arrs["status"] = st
    return arrs


def solve_fixed_grid(cfg: Config, params, tcoeffs, grid, init_std=0.0):
    params, tcoeffs, grid = _f64(params), _f64(tcoeffs), _f64(grid)
    K = grid.size
    arrs, out = _alloc(cfg, K)
    c = cfg.c()
    lib().pdqo_solve_fixed_grid(ctypes.byref(c), _p(params), _p(tcoeffs), _s(init_std), _p(grid), K, ctypes.byref(out))
    return _finish(cfg, arrs, out)


def lml(cfg: Config, params, tcoeffs, grid, every, data, stds, average=True, init_std=0.0):
    params, tcoeffs, grid, data, stds = map(_f64, (params, tcoeffs, grid, data, stds))
    stds2 = stds.reshape(data.shape[0], -1)
    c = cfg.c()
    return lib().pdqo_lml(ctypes.byref(c), _p(params), _p(tcoeffs), _s(init_std), _p(grid), grid.size, every, _p(data), _p(stds2), stds2.shape[1], int(average))


def lml_from_y0(cfg: Config, params, y0, grid, every, data, stds, average=True):
    params, y0, grid, data, stds = map(_f64, (params, y0, grid, data, stds))
    stds2 = stds.reshape(data.shape[0], -1)
    c = cfg.c()
    return _full(lib().pdqo_lml_from_y0(ctypes.byref(c), _p(params), _p(y0), _p(grid), grid.size, every, _p(data), _p(stds2), stds2.shape[1], int(average)))


def lml_from_y0_tape(cfg: Config, params, y0, grid, every, data, stds, tape, record: bool, average=True):
    """lml_from_y0 with the calibration tape: record=True stores the output scale of every
    solver_dynamic step into `tape` (an array of the active build's dtype), record=False replays
    them, so that the scales are constants (the reference's stop_gradient on the calibration)."""
    params, y0, grid, data, stds = map(_f64, (params, y0, grid, data, stds))
    stds2 = stds.reshape(data.shape[0], -1)
    assert tape.dtype == _dtype() and tape.flags.c_contiguous
    c = cfg.c()
    v = _full(lib().pdqo_lml_from_y0_tape(ctypes.byref(c), _p(params), _p(y0), _p(grid), grid.size, every, _p(data), _p(stds2), stds2.shape[1], int(average), _p(tape), tape.size, 1 if record else 2))
    if np.isnan(v):
        raise ValueError("calibration tape too short")
    return v


def sample(cfg: Config, params, sol: dict, base):
    """MarkovSequence.sample (estimators_and_losses.py:233-273) of a smoother solution `sol` (solve()'s
    outputs: term_mean/term_chol and the backward conditionals) with the given standard normal draws
    base (K, n).  Returns the K samples (K, mean_size)."""
    params, base = _f64(params), _f64(base)
    K = base.shape[0]
    arrs = [_f64(sol[k]) for k in ("term_mean", "term_chol", "cond_A", "cond_mean", "cond_chol", "cond_tl", "cond_to")]
    out = _zeros((K, arrs[0].size))
    c = cfg.c()
    rc = lib().pdqo_sample(ctypes.byref(c), _p(params), K, *[_p(a) for a in arrs], _p(base), _p(out))
    if rc:
        raise NotImplementedError("sampling is implemented for the dense and isotropic factorisations")
    return out


def step_error_power(cfg: Config, params, mean, chol, t, dt):
    """One step attempt from (mean, chol) at t and its error power (the configured error estimator).
    Returns (error_power, proposed mean)."""
    params, mean, chol = _f64(params), _f64(mean), _f64(chol)
    prop = _zeros(mean.shape)
    c = cfg.c()
    ep = lib().pdqo_step_error_power(ctypes.byref(c), _p(params), _p(mean), _p(chol), _s(t), _s(dt), _p(prop))
    return ep, prop


def lml_terminal(cfg: Config, params, mean, chol, u, std, tcoeff_index=0):
    """loss_lml_terminal_values (estimators_and_losses.py:20-50) of one marginal (mean, chol in
    solve()'s layouts) and data u (d,) with observation noise std (one entry, or one per coordinate
    for dense / blockdiag)."""
    params, mean, chol, u = _f64(params), _f64(mean), _f64(chol), _f64(u)
    std = _f64(np.atleast_1d(std))
    c = cfg.c()
    return lib().pdqo_lml_terminal(ctypes.byref(c), _p(params), _p(mean), _p(chol), _p(u), _p(std), std.size, int(tcoeff_index))


def offgrid_marginal(cfg: Config, params, t0, t1, t, output_scale, m0, L0, m1=None, L1=None):
    """offgrid_marginals (solvers.py:149-203) at one time t in (t0, t1): the filter predicts (m0, L0)
    from t0; the fixed-interval smoother reverts the filtering marginal (m0, L0) at t0 to t and t1 and
    marginalises the smoothed marginal (m1, L1) at t1 back to t.  Returns (mean, chol) in the
    layouts of solve()'s outputs."""
    params = _f64(params)
    m0, L0 = _f64(m0), _f64(L0)
    m1 = _f64(m0 if m1 is None else m1)
    L1 = _f64(L0 if L1 is None else L1)
    os_ = _f64(np.atleast_1d(output_scale))
    mo, Lo = _zeros(m0.shape), _zeros(L0.shape)
    c = cfg.c()
    rc = lib().pdqo_offgrid_marginal(ctypes.byref(c), _p(params), _s(t0), _s(t1), _s(t), _p(os_), _p(m0), _p(L0), _p(m1), _p(L1), _p(mo), _p(Lo))
    if rc == 2:
        raise NotImplementedError("offgrid marginals need a filter or a fixed-interval smoother")
    return mo, Lo


def lml_grad_fd(cfg: Config, params, y0, grid, every, data, stds, h=1e-6, average=True):
    params, y0, grid, data, stds = map(_f64, (params, y0, grid, data, stds))
    stds2 = stds.reshape(data.shape[0], -1)
    grad = _zeros(params.size + y0.size)
    c = cfg.c()
    val = lib().pdqo_lml_grad_fd(ctypes.byref(c), _p(params), params.size, _p(y0), _p(grid), grid.size, every, _p(data), _p(stds2), stds2.shape[1], int(average), _s(h), _p(grad))
    return val, grad


def solve_save_at_batch(cfg: Config, params, tcoeffs, save_at, n_threads=0):
    params, tcoeffs, save_at = _f64(params), _f64(tcoeffs), _f64(save_at)
    n_inst = params.shape[0]
    K = save_at.shape[1]
    term = _zeros((n_inst, cfg.mean_size))
    att = np.zeros(n_inst, dtype=np.int64)
    acc = np.zeros(n_inst, dtype=np.int64)
    c = cfg.c()
    st = lib().pdqo_solve_save_at_batch(ctypes.byref(c), n_inst, _p(params), params.shape[1], _p(tcoeffs), _p(save_at), K, _p(term), att.ctypes.data_as(_PL), acc.ctypes.data_as(_PL), n_threads)
    return {"term_mean": term, "attempts": att, "accepted": acc, "status": st}


def solve_save_every_step(cfg: Config, params, tcoeffs, t0, t1, init_std=0.0, K_max=20000):
    params, tcoeffs = _f64(params), _f64(tcoeffs)
    arrs, out = _alloc(cfg, K_max)
    c = cfg.c()
    K = lib().pdqo_solve_save_every_step(ctypes.byref(c), _p(params), _p(tcoeffs), _s(init_std), _s(t0), _s(t1), K_max, ctypes.byref(out))
    if K < 0:
        raise RuntimeError("K_max too small")
    arrs = _finish(cfg, arrs, out)
    for k, v in list(arrs.items()):
        if isinstance(v, np.ndarray) and k not in ("term_mean", "term_chol"):
            rows = K if v.shape[0] == K_max else (K - 1 if v.shape[0] == K_max - 1 else v.shape[0])
            arrs[k] = v[:rows]
    return arrs


def solve_fixed_grid_checkpoints(cfg: Config, params, tcoeffs, grid, every, init_std=0.0):
    params, tcoeffs, grid = _f64(params), _f64(tcoeffs), _f64(grid)
    K = (grid.size - 1) // every + 1
    arrs, out = _alloc(cfg, K)
    c = cfg.c()
    lib().pdqo_solve_fixed_grid_checkpoints(ctypes.byref(c), _p(params), _p(tcoeffs), _s(init_std), _p(grid), grid.size, every, ctypes.byref(out))
    return _finish(cfg, arrs, out)
