"""The solver's checks and their warnings (symhip_optimizer_warnings), against the oracle's
statuses and what each case makes true by construction:

* "Test optimizing with symbolic zeros on the diagonal" (test/symforce_optimizer_test.cc:
  432-459), written with a registered body: one Barron factor pulls x[0] to 3 while x[1:] have
  zero Jacobian columns, so their Hessian diagonal is zero before damping; with the default unit
  damping the optimizer converges to (3, 0, 0, 0, 0) with SUCCESS and warns of nothing;
* the same problem without damping leaves those entries zero after damping: CheckHessianDiagonal
  (levenberg_marquardt_solver.tcc:56-90) names indices 1-4, the singular solve gives a
  non-finite update (reported under debug_checks) and error, and the run fails with
  LAMBDA_OUT_OF_BOUNDS after the oracle's number of records;
* a factor whose convexity parameter makes alpha infinite linearizes to NaN: under debug_checks
  the warning names exactly that factor and its keys (opt/internal/linearizer_utils.h:307-367),
  and the run fails with INITIAL_ERROR_NOT_FINITE as the oracle's does."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from problem_builders import ProblemBuilder, INVALID
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu
EPS = 10 * np.finfo(np.float64).eps


def barron_problem(mus, ys, x0):
    b = ProblemBuilder()
    x = b.key("x", type_=capi.TYPE_VECTOR5, value=x0)
    e = b.key("e", value=EPS)
    yk = [b.key("y", i, INVALID, capi.TYPE_VECTOR5, y) for i, y in enumerate(ys)]
    mk = [b.key("u", i, INVALID, capi.TYPE_SCALAR, m) for i, m in enumerate(mus)]
    for k, m in zip(yk, mk):
        b.factor(capi.FACTOR_BARRON_VECTOR5, [x, k, m, e])
    return b.build()


def run(prob, params):
    opt = capi.HipOptimizer(prob, params)
    res = opt.optimize()
    return res, opt.warnings()


def test_symbolic_zeros_on_the_diagonal():
    prob = barron_problem([0.0], [[3, 0, 0, 0, 0]], np.zeros(5))
    res, warns = run(prob, capi.default_params())
    assert res["status"] == capi.STATUS_SUCCESS
    np.testing.assert_allclose(res["values"][:5], [3, 0, 0, 0, 0], atol=1e-6)
    assert warns == []
    ora = ow.OracleProblem(prob).optimize(capi.default_params())
    assert_same_trajectory(res, ora, rtol=1e-10)


@pytest.mark.parametrize("debug_checks", [0, 1])
def test_zero_diagonal_after_damping_is_reported(debug_checks):
    prob = barron_problem([0.0], [[3, 0, 0, 0, 0]], np.zeros(5))
    params = capi.default_params(use_unit_damping=0, use_diagonal_damping=0,
                                 debug_checks=debug_checks)
    res, warns = run(prob, params)
    ora = ow.OracleProblem(prob).optimize(params)
    assert (res["status"], res["failure_reason"]) == (ora["status"], ora["failure_reason"])
    assert res["failure_reason"] == capi.FAILURE_LAMBDA_OUT_OF_BOUNDS
    assert len(res["iterations"]) == len(ora["iterations"])
    zero = [w for w in warns if "Zero on diagonal after damping" in w]
    assert len(zero) == len(res["iterations"]) - 1  # once per damped factorization
    assert all(w.endswith("at indices: [1, 2, 3, 4]") for w in zero)
    assert any("Encountered non-finite error" in w for w in warns)
    upd = [w for w in warns if "Non-finite update" in w]
    assert bool(upd) == bool(debug_checks)
    assert any("Turn on debug_checks" in w for w in warns) == (not debug_checks)


@pytest.mark.parametrize("debug_checks", [0, 1])
def test_non_finite_factor_is_named(debug_checks):
    # alpha = 2 - 1 / (1 - mu + eps) is infinite at mu = 1 + eps
    prob = barron_problem([0.0, 1.0 + EPS, 0.0],
                          [[3, 0, 0, 0, 0], [1, 1, 1, 1, 1], [0, 0, 0, 0, 0]], np.zeros(5))
    params = capi.default_params(debug_checks=debug_checks)
    res, warns = run(prob, params)
    ora = ow.OracleProblem(prob).optimize(params)
    assert (res["status"], res["failure_reason"]) == (ora["status"], ora["failure_reason"])
    assert res["failure_reason"] == capi.FAILURE_INITIAL_ERROR_NOT_FINITE
    assert len(res["iterations"]) == 1
    assert any("Encountered non-finite initial error" in w for w in warns)
    named = [w for w in warns if "Non-finite linearization for factor" in w]
    if debug_checks:
        assert len(named) == 1
        w = named[0]
        assert w.startswith("LM<symhip> Non-finite linearization for factor 1 (type 11)")
        for key in ("x (offset=0, size=5)", "y_1 (offset=11, size=5)",
                    "u_1 (offset=22, size=1)", "e (offset=5, size=1)"):
            assert key in w, (key, w)
        assert "Residual: [nan" in w
    else:
        assert named == []
