"""Iterative form of the Schur solve: block-Jacobi PCG on the reduced camera system (no
reference counterpart; SparseSchurSolver factorizes S directly, sparse_schur_solver.tcc:133-163).

Checked against the direct path it replaces: at a tight tolerance the LM trajectory is the
direct one; the fallback reproduces the direct solve exactly; at a loose tolerance the optimizer
still reaches the direct path's final cost."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu


def run(prob, params, **opts):
    opt = capi.HipOptimizer(prob, params, capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                                                              **opts))
    res = opt.optimize()
    return res, opt.linear_solver_stats()


@pytest.mark.parametrize("gen", [lambda: capi.gen_ba_snavely(20, 600, 11),
                                 lambda: capi.gen_ba_circle(10, 300, 4)])
def test_tight_pcg_follows_direct_trajectory(gen):
    prob = gen()
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=25)
    direct, _ = run(prob, params)
    pcg, st = run(prob, params, use_pcg=2, pcg_max_iters=2000, pcg_rel_tol=1e-13)
    assert st["solves"] > 0 and st["fallbacks"] == 0
    # a 1e-13 relative residual moves the step by ~cond(S) * 1e-13
    assert_same_trajectory(pcg, direct, rtol=1e-6)


def test_fallback_is_the_direct_solve():
    prob = capi.gen_ba_snavely(15, 400, 21)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=15)
    direct, _ = run(prob, params)
    # one CG iteration never reaches 1e-12: in strict mode every solve falls back to LDL^T
    fb, st = run(prob, params, use_pcg=2, pcg_max_iters=1, pcg_rel_tol=1e-12)
    assert st["fallbacks"] == st["solves"] > 0
    assert_same_trajectory(fb, direct, rtol=1e-12)
    np.testing.assert_array_equal(fb["values"], direct["values"])


def test_loose_pcg_reaches_direct_final_cost():
    prob = capi.gen_ba_snavely(30, 1500, 7)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100)
    direct, _ = run(prob, params)
    pcg, st = run(prob, params, use_pcg=1, pcg_max_iters=300, pcg_rel_tol=1e-4)
    best = lambda r: min(it["new_error"] for it in r["iterations"])
    assert best(pcg) <= best(direct) * (1 + 1e-6)
    assert st["fallbacks"] == 0 and st["max_iterations"] <= 300


def test_truncated_pcg_keeps_inexact_steps():
    prob = capi.gen_ba_snavely(30, 1500, 7)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100)
    direct, _ = run(prob, params)
    pcg, st = run(prob, params, use_pcg=1, pcg_max_iters=10, pcg_rel_tol=1e-8)
    # every solve stops at 10 CG iterations and its iterate is used (no LDL^T fallback)
    assert st["fallbacks"] == 0 and st["max_iterations"] == 10
    assert st["iterations"] == 10 * st["solves"]
    best = lambda r: min(it["new_error"] for it in r["iterations"])
    assert best(pcg) < 0.02 * pcg["iterations"][0]["new_error"]  # LM still converges
    assert best(pcg) <= best(direct) * (1 + 1e-2)


def test_switch_to_exact_solves_reaches_direct_cost():
    prob = capi.gen_ba_snavely(40, 3000, 13)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100)
    direct, _ = run(prob, params)
    fast, st = run(prob, params, use_pcg=1, pcg_max_iters=20, pcg_rel_tol=1e-2,
                   pcg_exact_below_reduction=1e-3)
    best = lambda r: min(it["new_error"] for it in r["iterations"])
    assert st["solves"] < len(fast["iterations"]) - 1  # the late iterations ran the LDL^T
    assert best(fast) <= best(direct) * (1 + 1e-6)


def test_forcing_tightens_pcg_and_reaches_direct_minimum():
    """Inexact Newton: after an accepted step with relative reduction r the next PCG solve runs
    to min(tol, c sqrt(r)) with a proportionally larger iteration limit. Alone it leaves LM's
    early exit a few 1e-6 above the minimum (its last steps are still inexact); with a late
    switch to the exact LDL^T it converges to the direct path's minimum."""
    prob = capi.gen_ba_snavely(40, 3000, 13)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100,
                                 early_exit_min_reduction=0.0)
    direct, _ = run(prob, params)
    forced, st = run(prob, params, use_pcg=1, pcg_max_iters=20, pcg_rel_tol=1e-2,
                     pcg_forcing=0.3, pcg_exact_below_reduction=1e-5)
    best = lambda r: min(it["new_error"] for it in r["iterations"])
    assert st["fallbacks"] == 0 and 0 < st["solves"] < len(forced["iterations"]) - 1
    assert st["max_iterations"] > 20  # the tightened solves ran past the base limit
    assert abs(best(forced) - best(direct)) <= 1e-8 * best(direct)


def test_pcg_rejected_on_cholesky_path():
    prob = capi.gen_pose_smoothing(10, 42)
    with pytest.raises(RuntimeError, match="Schur path only"):
        capi.HipOptimizer(prob, capi.default_params(), capi.solver_options(use_pcg=1))
