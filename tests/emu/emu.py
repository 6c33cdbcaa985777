"""Loader of the CPU build of the kernels (test infrastructure)."""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libpdqemu.so")


def build(extra=()) -> str:
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "pdq_emu.cpp")
    # compiled to a private name and renamed into place, so that concurrent
    # test processes never load a half-written library
    tmp = f"{LIB}.{os.getpid()}.tmp"
    cmd = ["g++", "-O1", "-std=c++20", "-pthread", "-fPIC", "-shared", "-DPDQ_HOST_EMU", *extra, "-o", tmp, src]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def solve(cfg, params, tcoeffs, times, fixed_grid=False, kernel="coop"):
    """cfg: probdiffeq_amd.capi.Config; numpy inputs; returns numpy outputs.

    kernel: "coop" (the cooperative kernel's code) or "lane" (the
    lane-per-instance isotropic kernel's code)."""
    from probdiffeq_amd import capi

    params = np.ascontiguousarray(params, dtype=np.float64)
    tcoeffs = np.ascontiguousarray(tcoeffs, dtype=np.float64)
    times = np.ascontiguousarray(times, dtype=np.float64)
    B, K = times.shape
    ms, cs, nn, nv = cfg.mean_size, cfg.chol_size, cfg.nblk * cfg.n * cfg.n, cfg.nblk * cfg.n
    out = {
        "t": np.zeros((B, K)),
        "mean": np.zeros((B, K, ms)),
        "chol": np.zeros((B, K, cs)),
        "output_scale": np.zeros((B, K, cfg.nblk)),
        "num_steps": np.zeros((B, K - 1), dtype=np.int64),
        "attempts": np.zeros(B, dtype=np.int64),
        "accepted": np.zeros(B, dtype=np.int64),
        "status": np.zeros(B, dtype=np.int32),
    }
    if cfg.strategy != "filter":
        out.update(
            filt_mean=np.zeros((B, K, ms)),
            filt_chol=np.zeros((B, K, cs)),
            cond_A=np.zeros((B, K - 1, nn)),
            cond_mean=np.zeros((B, K - 1, ms)),
            cond_chol=np.zeros((B, K - 1, cs)),
            cond_tl=np.zeros((B, K - 1, nv)),
            cond_to=np.zeros((B, K - 1, nv)),
            term_mean=np.zeros((B, ms)),
            term_chol=np.zeros((B, cs)),
        )
    s = capi._Solution()
    for name, _ in capi._Solution._fields_:
        a = out.get(name)
        setattr(s, name, a.ctypes.data if a is not None else None)
    c = cfg.c(host=True)
    P = ctypes.c_void_p
    fn = lib().pdqemu_solve_lane if kernel == "lane" else lib().pdqemu_solve
    rc = fn(ctypes.byref(c), B, P(params.ctypes.data), P(tcoeffs.ctypes.data), P(times.ctypes.data), K, ctypes.byref(s), int(fixed_grid))
    if rc:
        raise RuntimeError(f"emulation rc {rc}")
    return out


def lml(cfg, params, y0, grid, every, data, stds, average=True, grad=True, kernel="coop", tangents_per_lane=None):
    """kernel: "coop" (the cooperative kernel's LML) or "lane" (the lane-per-instance LML with
    `tangents_per_lane` tangent columns per lane; default 1, the device library's choice)."""
    params, y0, grid, data, stds = (np.ascontiguousarray(x, dtype=np.float64) for x in (params, y0, grid, data, stds))
    B = params.shape[0]
    value = np.zeros(B)
    g = np.zeros((B, params.shape[1] + y0.shape[1])) if grad else None
    n_std = stds.shape[-1] if stds.ndim == 3 else 1
    c = cfg.c(host=True)
    P = ctypes.c_void_p
    if kernel == "lane":
        n_tan = params.shape[1] + y0.shape[1] if grad else 0
        tpl = tangents_per_lane if tangents_per_lane is not None else (0 if n_tan == 0 else 1)
        rc = lib().pdqemu_lml_lane(ctypes.byref(c), B, P(params.ctypes.data), P(y0.ctypes.data), P(grid.ctypes.data), grid.size, every, P(data.ctypes.data), P(stds.ctypes.data), int(average), P(value.ctypes.data), P(g.ctypes.data) if grad else None, tpl)
    else:
        rc = lib().pdqemu_lml(ctypes.byref(c), B, P(params.ctypes.data), P(y0.ctypes.data), P(grid.ctypes.data), grid.size, every, P(data.ctypes.data), P(stds.ctypes.data), n_std, int(average), P(value.ctypes.data), P(g.ctypes.data) if grad else None)
    if rc:
        raise RuntimeError(f"emulation rc {rc}")
    return value, g


def lml_adjoint(cfg, params, y0, grid, every, data, stds, average=True, grad=True, lanes=3):
    """The reverse-mode adjoint LML kernel's code (pdq_lane_adjoint.h) on the CPU; `lanes` scratch
    columns shared by the instances in grid-stride order, as on the device."""
    params, y0, grid, data, stds = (np.ascontiguousarray(x, dtype=np.float64) for x in (params, y0, grid, data, stds))
    B = params.shape[0]
    value = np.zeros(B)
    g = np.zeros((B, params.shape[1] + y0.shape[1])) if grad else None
    c = cfg.c(host=True)
    P = ctypes.c_void_p
    rc = lib().pdqemu_lml_adjoint(ctypes.byref(c), B, P(params.ctypes.data), P(y0.ctypes.data), P(grid.ctypes.data), grid.size, every, P(data.ctypes.data), P(stds.ctypes.data), int(average), P(value.ctypes.data), P(g.ctypes.data) if grad else None, lanes)
    if rc:
        raise RuntimeError(f"emulation rc {rc}")
    return value, g


def offgrid(cfg, times, sol, t_query):
    """The off-grid marginal kernel's code on the CPU: times (B, K), sol = dict with mean, chol,
    output_scale (B, K, 1) and, for the fixed-interval smoother, filt_mean / filt_chol, in
    solve()'s layouts; t_query (B, Q).  Returns (mean (B, Q, N d), chol (B, Q, N N))."""
    arr = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    times, t_query = arr(times), arr(t_query)
    B, K = times.shape
    Qn = t_query.shape[1]
    N, d = cfg.num_derivatives + 1, cfg.ode_dim
    mean_out = np.zeros((B, Qn, N * d))
    chol_out = np.zeros((B, Qn, cfg.chol_size))
    m, L, os_ = arr(sol["mean"]), arr(sol["chol"]), arr(sol["output_scale"])
    fm = arr(sol["filt_mean"]) if "filt_mean" in sol else None
    fL = arr(sol["filt_chol"]) if "filt_chol" in sol else None
    P = ctypes.c_void_p
    ptr = lambda x: P(x.ctypes.data) if x is not None else None
    c = cfg.c(host=True)
    rc = lib().pdqemu_offgrid(ctypes.byref(c), B, K, ptr(times), ptr(m), ptr(L), ptr(fm), ptr(fL), ptr(os_), Qn, ptr(t_query), ptr(mean_out), ptr(chol_out))
    if rc:
        raise RuntimeError(f"emulation rc {rc}")
    return mean_out, chol_out


def sample(cfg, sol, base):
    """The posterior sampler's code on the CPU: sol = one instance's term_*/cond_* (solve()'s layouts),
    base (S, K, n) with n the covariance dimension.  Returns (S, K, mean_size)."""
    arr = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    base = arr(base)
    S, K, _ = base.shape
    out = np.zeros((S, K, cfg.mean_size))
    keys = ("term_mean", "term_chol", "cond_A", "cond_mean", "cond_chol", "cond_tl", "cond_to")
    arrs = [arr(sol[k]) for k in keys]
    P = ctypes.c_void_p
    c = cfg.c(host=True)
    rc = lib().pdqemu_sample(ctypes.byref(c), 1, K, *[P(a.ctypes.data) for a in arrs], S, P(base.ctypes.data), P(out.ctypes.data))
    if rc:
        raise RuntimeError(f"emulation rc {rc}")
    return out
