"""GPU: a captured solve graph replays to the eager results, also after its inputs change in place."""

import numpy as np
import pytest

import test_gpu_parity as T

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _inputs(seed):
    from probdiffeq_amd import capi

    params, y0, times = T._lv_batch(64, seed=seed)
    P, Y, S = (torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda() for x in (params, y0, times))
    return P, capi.taylor_coefficients("lotka_volterra", P, Y, 0.0, 4), S


def test_solve_graph_replays_match_eager_solves():
    from probdiffeq_amd import capi, graphs

    capi.require_device()
    cfg = capi.Config(factorisation="isotropic", num_derivatives=4, strategy="fixedpoint", solver="dynamic")
    P, tc, S = _inputs(1)
    g = graphs.SolveGraph(cfg, P, tc, S)
    keys = ("t", "mean", "chol", "output_scale", "filt_mean", "cond_A", "cond_chol", "attempts", "status")
    for replay in range(3):
        eager = capi.solve(cfg, P, tc, S)
        torch.cuda.synchronize()
        out = g.replay()
        torch.cuda.synchronize()
        for k in keys:
            assert torch.equal(out[k], eager[k]), (replay, k)
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
    # new inputs written in place: the replay solves the new problem
    P2, tc2, S2 = _inputs(2)
    P.copy_(P2)
    tc.copy_(tc2)
    S.copy_(S2)
    out = g.replay()
    eager = capi.solve(cfg, P2, tc2, S2)
    torch.cuda.synchronize()
    for k in keys:
        assert torch.equal(out[k], eager[k]), k


@pytest.mark.parametrize(
    "kw",
    [
        dict(factorisation="isotropic", num_derivatives=4, strategy="fixedpoint", lanes_per_instance=16),
        dict(factorisation="dense", num_derivatives=4, strategy="filter"),
        dict(factorisation="dense", num_derivatives=3, strategy="fixedinterval", solver="mle"),
        dict(factorisation="blockdiag", num_derivatives=3, strategy="fixedpoint"),
    ],
    ids=["isotropic-coop16", "dense-filter", "dense-fixedinterval-mle", "blockdiag"],
)
def test_solve_graph_captures_the_cooperative_and_blockdiag_kernels(kw):
    """save_at graphs go through the no-allocation entry point: its matrices and work-queue counter
    are written by kernels into the fixed workspace, so the cooperative and blockdiag solves capture
    too.  Every replay equals the eager solve, with eager work between replays, also after the
    inputs change in place."""
    from probdiffeq_amd import capi, graphs

    capi.require_device()
    cfg = capi.Config(**kw)
    n = 16
    params, y0, times = T._lv_batch(n, seed=5)
    P, Y, S = (torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda() for x in (params, y0, times))
    tc = capi.taylor_coefficients("lotka_volterra", P, Y, 0.0, cfg.num_derivatives)
    g = graphs.SolveGraph(cfg, P, tc, S)
    keys = [k for k in ("t", "mean", "chol", "output_scale", "filt_mean", "cond_A", "cond_chol", "attempts", "status") if k in g.out]
    for replay in range(3):
        print(f"graph replay {replay} {kw}", flush=True)
        eager = capi.solve(cfg, P, tc, S)
        torch.cuda.synchronize()
        out = g.replay()
        torch.cuda.synchronize()
        for k in keys:
            assert torch.equal(out[k], eager[k]), (replay, k)
        assert int(out["status"].abs().sum()) == 0
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
    params2, y02, times2 = T._lv_batch(n, seed=6)
    P.copy_(torch.as_tensor(params2, dtype=torch.float64))
    S.copy_(torch.as_tensor(times2, dtype=torch.float64))
    tc.copy_(capi.taylor_coefficients("lotka_volterra", P, torch.as_tensor(y02, dtype=torch.float64).cuda(), 0.0, cfg.num_derivatives))
    out = g.replay()
    eager = capi.solve(cfg, P, tc, S)
    torch.cuda.synchronize()
    for k in keys:
        assert torch.equal(out[k], eager[k]), k


def test_solve_graph_refuses_cooperative_fixed_grids():
    from probdiffeq_amd import capi, graphs

    cfg = capi.Config(factorisation="isotropic", num_derivatives=4, lanes_per_instance=16)
    P, tc, S = _inputs(3)
    with pytest.raises(NotImplementedError):
        graphs.SolveGraph(cfg, P, tc, S, fixed_grid=True)
