"""Schur backend (landmark elimination + RCS + supernodal LDL^T) against the oracle's
SparseSchurSolver restatement (sparse_schur_solver.tcc:16-163) driven by the same LM loop."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,gen,lam", [
    ("ba_snavely", lambda: capi.gen_ba_snavely(12, 300, 5), capi.LAMBDA_DYNAMIC),
    ("ba_snavely_static", lambda: capi.gen_ba_snavely(20, 500, 9), capi.LAMBDA_STATIC),
    ("ba_circle", lambda: capi.gen_ba_circle(8, 200, 3), capi.LAMBDA_STATIC),
])
@pytest.mark.parametrize("ordering", [capi.ORDERING_NATURAL, capi.ORDERING_MIN_DEGREE])
def test_schur_matches_oracle(name, gen, lam, ordering):
    prob = gen()
    c_dim = int(prob.info["c_dim"])
    params = capi.default_params(lambda_update_type=lam, iterations=30)
    hip = capi.HipOptimizer(prob, params, capi.solver_options(
        linear_solver=capi.SOLVER_SCHUR, ordering=ordering)).optimize()
    ora = ow.OracleProblem(prob).optimize(params, solver=ow.SCHUR, c_dim=c_dim)
    assert_same_trajectory(hip, ora, rtol=1e-8)
    np.testing.assert_allclose(hip["values"], ora["values"], rtol=0,
                               atol=1e-7 * max(1.0, np.abs(ora["values"]).max()))


def test_schur_and_cholesky_paths_agree():
    prob = capi.gen_ba_snavely(15, 400, 21)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=20)
    a = capi.HipOptimizer(prob, params, capi.solver_options(linear_solver=capi.SOLVER_SCHUR)).optimize()
    b = capi.HipOptimizer(prob, params, capi.solver_options(linear_solver=capi.SOLVER_CHOLESKY)).optimize()
    assert_same_trajectory(a, b, rtol=1e-8)
