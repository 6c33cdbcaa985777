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
