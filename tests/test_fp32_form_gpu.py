"""FP32 precision form of the Schur path (SYMHIP_PRECISION_F32_JAC): the Hessian-side products
come from the FP32 factor bodies, W, Z and the PCG copy of S are stored in FP32, while residual,
gradient, the matrix-core sums, LDL^T, cost and gain ratio stay FP64; once a PCG run switches to
the exact LDL^T it also returns to FP64 storage (no reference counterpart: the reference runs
the whole SparseSchurSolver in the optimizer's Scalar, opt/sparse_schur_solver.tcc:16-163).

The bar is the FP64 method's: the same initial cost bit for bit (cost is FP64), steps that
differ from the FP64 steps only by FP32 rounding amplified by the cancellation in
S = B - sum Z Z^T (measured: about 1e-3 of a step's cost reduction), and the same minimum: with
early exit disabled both runs converge, and the FP64 gradient fixes the point they converge to,
so the best costs agree to 1e-8 relative."""
import pytest

from symforce_amd import _capi as capi

pytestmark = pytest.mark.gpu


def run(prob, params, **opts):
    opt = capi.HipOptimizer(prob, params, capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                                                              **opts))
    return opt.optimize(), opt.linear_solver_stats()


def best(r):
    return min(it["new_error"] for it in r["iterations"])


def converge_params(iterations):
    return capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=iterations,
                               early_exit_min_reduction=0.0)


@pytest.mark.parametrize("gen", [lambda: capi.gen_ba_snavely(30, 1500, 7),
                                 lambda: capi.gen_ba_circle(10, 300, 4)])
def test_fp32_direct_tracks_fp64(gen):
    prob = gen()
    params = converge_params(150)
    f64, _ = run(prob, params)
    f32, _ = run(prob, params, precision=capi.PRECISION_F32_JAC)
    it64, it32 = f64["iterations"], f32["iterations"]
    # the initial cost (first record) is evaluated in FP64 on the same Values
    assert it32[0]["new_error"] == it64[0]["new_error"]
    # the first step's cost differs only by the FP32 rounding of the Hessian blocks
    e64, e32 = it64[1]["new_error"], it32[1]["new_error"]
    assert abs(e32 - e64) <= 1e-2 * abs(it64[0]["new_error"] - e64)
    assert abs(best(f32) - best(f64)) <= 1e-8 * best(f64)


def test_fp32_pcg_with_exact_switch_reaches_fp64_minimum():
    prob = capi.gen_ba_snavely(40, 3000, 13)
    params = converge_params(100)
    f64, _ = run(prob, params)
    fast, st = run(prob, params, precision=capi.PRECISION_F32_JAC, use_pcg=1, pcg_max_iters=20,
                   pcg_rel_tol=1e-2, pcg_exact_below_reduction=1e-3)
    assert st["solves"] > 0 and st["fallbacks"] == 0
    assert st["solves"] < len(fast["iterations"]) - 1  # the late iterations ran the FP64 LDL^T
    assert abs(best(fast) - best(f64)) <= 1e-8 * best(f64)


def test_fp32_pcg_strict_tight_follows_fp64_direct():
    prob = capi.gen_ba_snavely(20, 600, 11)
    params = converge_params(100)
    f64, _ = run(prob, params)
    f32, st = run(prob, params, precision=capi.PRECISION_F32_JAC, use_pcg=2, pcg_max_iters=2000,
                  pcg_rel_tol=1e-10)
    assert st["fallbacks"] == 0
    # FP32-stored S: each step's cost agrees with the FP64 direct step's to FP32 accuracy
    for a, b in zip(f32["iterations"][1:6], f64["iterations"][1:6]):
        assert abs(a["new_error"] - b["new_error"]) <= 1e-4 * b["new_error"]
    assert abs(best(f32) - best(f64)) <= 1e-8 * best(f64)


def test_fp32_form_scope_is_checked():
    prob = capi.gen_ba_snavely(10, 200, 3)
    params = capi.default_params(iterations=5)
    with pytest.raises(RuntimeError, match="Schur path only"):
        capi.HipOptimizer(prob, params, capi.solver_options(precision=capi.PRECISION_F32_JAC))
    with pytest.raises(RuntimeError, match="FP64 systems"):
        capi.HipOptimizer(prob, params, capi.solver_options(
            linear_solver=capi.SOLVER_SCHUR, precision=capi.PRECISION_F32_JAC, use_pcg=1,
            pcg_preconditioner=capi.PRECOND_LEARNED))
