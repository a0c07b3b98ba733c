"""The 1000-camera bundle adjustment at its full size (1000 cameras, 200 000 points, ~1M
observations; the metric's workload), held to a size-independent property: the oracle needs
~45 minutes per LM run there, so instead the exact Schur step (reduced camera system assembled on
the matrix cores, supernodal LDL^T of the 9000 x 9000 system with nested dissection: 38 levels,
both update-tile sizes, the fused forward solve) is compared with an independent solver of the
same reduced system, strict FP64 block-Jacobi PCG run to a relative residual of 1e-12 (strict
mode: no step may fall back to the LDL^T). The two steps differ by cond(S) times the PCG
residual, so the costs along the first LM iterations agree to 1e-6 relative (the bar the small
problems' tight-PCG test uses); a factorization error at this size (a wrong tile, a missed
update, a race between levels) moves a step by far more."""
import pytest

from symforce_amd import _capi as capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba1k():
    return capi.gen_ba_snavely(1000, 200000, 42)


def test_ba1k_exact_step_matches_tight_pcg(ba1k):
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=4,
                                 early_exit_min_reduction=0.0)
    base = dict(linear_solver=capi.SOLVER_SCHUR, ordering=capi.ORDERING_NESTED_DISSECTION)
    direct = capi.HipOptimizer(ba1k, params, capi.solver_options(**base)).optimize()
    opt = capi.HipOptimizer(ba1k, params, capi.solver_options(
        **base, use_pcg=2, pcg_max_iters=20000, pcg_rel_tol=1e-12))
    pcg = opt.optimize()
    st = opt.linear_solver_stats()
    assert st["solves"] == len(pcg["iterations"]) - 1 and st["fallbacks"] == 0
    assert len(direct["iterations"]) == len(pcg["iterations"]) == 5
    assert direct["iterations"][0]["new_error"] == pcg["iterations"][0]["new_error"]
    for d, p in zip(direct["iterations"][1:], pcg["iterations"][1:]):
        assert abs(d["new_error"] - p["new_error"]) <= 1e-6 * d["new_error"]
        assert d["update_accepted"] == p["update_accepted"]
