"""GPU: the lanes-per-instance autotuner records its choice and later calls use it."""

import numpy as np
import pytest

import test_gpu_parity as T

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def test_autotune_records_the_fastest_choice():
    from probdiffeq_amd import autotune, capi

    capi.require_device()
    params, y0, times = T._lv_batch(96, seed=4)
    P, Y, S = (torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda() for x in (params, y0, times))
    tc = capi.taylor_coefficients("lotka_volterra", P, Y, 0.0, 4)
    cfg = capi.Config(factorisation="isotropic", num_derivatives=4, strategy="fixedpoint", solver="dynamic")
    timings, best = autotune.tune_lanes(cfg, P, tc, S, candidates=(1, 16, 32), reps=1)
    print("autotune timings (ms):", timings, "best:", best, flush=True)
    assert set(timings) == {1, 16, 32} and all(v is not None and v > 0 for v in timings.values())
    assert best == min(timings, key=timings.get)
    import ctypes

    c = cfg.c()
    assert capi.lib().pdq_tuned_lanes(ctypes.byref(c), 96) == best
    out_auto = capi.solve(cfg, P, tc, S)
    cfg_best = capi.Config(factorisation="isotropic", num_derivatives=4, strategy="fixedpoint", solver="dynamic", lanes_per_instance=best)
    out_best = capi.solve(cfg_best, P, tc, S)
    torch.cuda.synchronize()
    for key in ("mean", "chol", "attempts", "cond_A"):
        assert torch.equal(out_auto[key], out_best[key]), key


def test_autotune_skips_choices_the_configuration_refuses():
    from probdiffeq_amd import autotune, capi

    params, y0, times = T._lv_batch(8, seed=2)
    P, Y, S = (torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda() for x in (params, y0, times))
    tc = capi.taylor_coefficients("lotka_volterra", P, Y, 0.0, 4)
    cfg = capi.Config(factorisation="dense", num_derivatives=4)
    timings, best = autotune.tune_lanes(cfg, P, tc.contiguous(), S, candidates=(1, 64), reps=1, record=False)
    assert timings[1] is None and timings[64] is not None and best == 64


def test_lml_autotune_records_the_fastest_cooperative_group():
    """tune_lml_lanes times the LML at each group size and records the fastest cooperative one:
    a later call with lanes_per_instance = 0 runs that group size (bitwise the same value and
    gradient as the explicit choice)."""
    import lml_cases as LC
    from probdiffeq_amd import autotune, capi

    capi.require_device()
    params, y0, grid, every, data, stds = LC.lml_inputs(24, seed=21, t_end=6.0, n_grid=121, every=10)
    dev = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda()  # noqa: E731
    args = (dev(params), dev(y0), dev(grid), every, dev(data), dev(stds))
    cfg = capi.Config(factorisation="dense", num_derivatives=3, ode_dim=2, strategy="fixedpoint", solver="dynamic")
    timings, best = autotune.tune_lml_lanes(cfg, *args, candidates=(1, 8, 32, 64), reps=1)
    print("lml autotune timings (ms):", timings, "best:", best, flush=True)
    assert timings[1] is None  # the lane LML is isotropic only
    assert all(timings[k] is not None and timings[k] > 0 for k in (8, 32, 64))
    assert best == min((k for k in timings if timings[k] is not None), key=timings.get)
    v0, g0 = capi.lml_value_and_grad(cfg, *args)
    vb, gb = capi.lml_value_and_grad(capi.Config(**{**cfg.__dict__, "lanes_per_instance": best}), *args)
    torch.cuda.synchronize()
    assert torch.equal(v0, vb) and torch.equal(g0, gb)
