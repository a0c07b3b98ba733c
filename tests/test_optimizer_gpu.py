"""Device LM optimizer (Cholesky path) against the reference's known answers and, iteration by
iteration, against the FP64 oracle on the same seeded inputs."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
import problem_builders as pb

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def run_both(prob, params, ordering=capi.ORDERING_NATURAL):
    hip = capi.HipOptimizer(prob, params, capi.solver_options(ordering=ordering)).optimize()
    ora = ow.OracleProblem(prob).optimize(params)
    return hip, ora


def assert_same_trajectory(hip, ora, rtol=RTOL, lambda_rtol=1e-8):
    assert hip["status"] == ora["status"]
    assert hip["failure_reason"] == ora["failure_reason"]
    assert len(hip["iterations"]) == len(ora["iterations"])
    assert hip["best_index"] == ora["best_index"]
    for a, b in zip(hip["iterations"], ora["iterations"]):
        assert a["iteration"] == b["iteration"]
        assert a["update_accepted"] == b["update_accepted"]
        # DYNAMIC lambda multiplies by a function of the gain ratio, a quotient of two small
        # differences, so ~1e-15 cost differences move lambda at ~1e-11 relative
        assert a["current_lambda"] == pytest.approx(b["current_lambda"], rel=lambda_rtol)
        for k in ("new_error", "new_error_linear"):
            assert a[k] == pytest.approx(b[k], rel=rtol, abs=1e-12), (k, a, b)


def test_pose_smoothing_known_answer():
    """test/symforce_optimizer_test.cc:142-176."""
    params = capi.default_params(iterations=50, early_exit_min_reduction=1e-4)
    hip, ora = run_both(capi.gen_pose_smoothing(10, 42), params)
    last = hip["iterations"][-1]
    assert last["iteration"] == 12
    assert last["current_lambda"] == pytest.approx(0.0039, rel=1e-1)
    assert last["new_error"] == pytest.approx(7.801, rel=1e-3)
    assert hip["status"] == capi.STATUS_SUCCESS
    assert_same_trajectory(hip, ora)


def test_pose_smoothing_min_degree_ordering_same_trajectory():
    params = capi.default_params(iterations=50, early_exit_min_reduction=1e-4)
    hip, ora = run_both(capi.gen_pose_smoothing(10, 42), params, capi.ORDERING_MIN_DEGREE)
    assert_same_trajectory(hip, ora)


def test_dynamic_lambda_27_iterations():
    """test/symforce_optimizer_test.cc:461-475."""
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC)
    hip, ora = run_both(capi.gen_pose_smoothing(10, 42), params)
    assert hip["status"] == capi.STATUS_SUCCESS
    assert len(hip["iterations"]) == 27
    assert_same_trajectory(hip, ora)


def test_best_linearization_shape():
    """test/symforce_optimizer_test.cc:477-506: residual 66, H 60x60 nnz 534, all finite."""
    prob = capi.gen_pose_smoothing(10, 42)
    opt = capi.HipOptimizer(prob, capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC))
    res = opt.optimize()
    lin = opt.best_linearization(66, 534, 60)
    for v in lin.values():
        assert np.isfinite(v).all()
    want = ow.OracleProblem(prob).linearize(res["values"])
    np.testing.assert_allclose(lin["residual"], want["residual"], rtol=1e-9, atol=1e-12)


def test_rot3_smoothing_known_answer():
    """test/symforce_optimizer_test.cc:183-264."""
    params = capi.default_params(iterations=50, early_exit_min_reduction=1e-4)
    hip, ora = run_both(capi.gen_rot_smoothing(10, 42), params)
    last = hip["iterations"][-1]
    assert last["iteration"] == 6
    assert last["current_lambda"] == pytest.approx(2.4e-4, rel=1e-1)
    assert last["new_error"] == pytest.approx(2.174, rel=1e-3)
    assert_same_trajectory(hip, ora)


def test_frozen_out_of_order_keys():
    """test/symforce_optimizer_test.cc:276-348."""
    params = capi.default_params(iterations=50, early_exit_min_reduction=1e-4)
    hip, ora = run_both(capi.gen_rot_frozen(3, 42), params)
    last = hip["iterations"][-1]
    assert last["iteration"] == 5
    assert last["current_lambda"] < 1e-3
    assert last["new_error"] < 1e-15
    assert hip["status"] == capi.STATUS_SUCCESS


def test_robot_3d_localization_known_answer():
    """test/symforce_examples_robot_3d_localization_test.py:49-51."""
    prob, d = pb.robot3d_problem()
    params = capi.default_params(initial_lambda=1e4, lambda_down_factor=0.5)
    hip, ora = run_both(prob, params)
    assert hip["iterations"][0]["new_error"] == pytest.approx(463700.5576620833, abs=5e-7)
    assert hip["iterations"][hip["best_index"]]["new_error"] < 140
    assert hip["status"] == capi.STATUS_SUCCESS
    assert_same_trajectory(hip, ora)


def test_status_hit_iteration_limit():
    params = capi.default_params(iterations=5)
    hip, ora = run_both(capi.gen_pose_smoothing(10, 42), params)
    assert hip["status"] == capi.STATUS_HIT_ITERATION_LIMIT
    assert hip["failure_reason"] == capi.FAILURE_INVALID
    assert_same_trajectory(hip, ora)


def test_status_lambda_out_of_bounds():
    """An exactly-solved problem never reduces the error; with lambda at its upper bound the step
    is rejected and LM fails with LAMBDA_OUT_OF_BOUNDS (levenberg_marquardt_solver.tcc:287-290)."""
    b = pb.ProblemBuilder()
    x = b.key("x", 0, type_=capi.TYPE_POSE3, value=[0, 0, 0, 1, 1, 2, 3])
    prior = b.key("a", type_=capi.TYPE_POSE3, value=[0, 0, 0, 1, 1, 2, 3])
    s = b.key("s", type_=capi.TYPE_MATRIX66, value=np.eye(6))
    e = b.key("e", value=1e-10)
    b.factor(capi.FACTOR_PRIOR_POSE3, [x, prior, s, e])
    params = capi.default_params(initial_lambda=1e6, lambda_upper_bound=1e6)
    hip, ora = run_both(b.build(), params)
    assert hip["status"] == capi.STATUS_FAILED
    assert hip["failure_reason"] == capi.FAILURE_LAMBDA_OUT_OF_BOUNDS
    assert_same_trajectory(hip, ora)


def test_status_initial_error_not_finite():
    prob = capi.gen_pose_smoothing(4, 1)
    prob.values[4] = np.nan
    params = capi.default_params()
    hip, ora = run_both(prob, params)
    assert hip["status"] == capi.STATUS_FAILED
    assert hip["failure_reason"] == capi.FAILURE_INITIAL_ERROR_NOT_FINITE
    assert ora["failure_reason"] == capi.FAILURE_INITIAL_ERROR_NOT_FINITE
    assert len(hip["iterations"]) == 1


@pytest.mark.parametrize("name,gen,params", [
    ("ba_circle", lambda: capi.gen_ba_circle(8, 150, 3), dict()),
    ("pose_graph", lambda: capi.gen_pose_graph(2000, 150, 7), dict()),
    ("localization", lambda: capi.gen_localization(30, 25, 0.3, 5), dict()),
    ("ba_snavely_cholesky", lambda: capi.gen_ba_snavely(10, 200, 5),
     dict(lambda_update_type=capi.LAMBDA_DYNAMIC)),
])
def test_generated_problems_match_oracle(name, gen, params):
    prob = gen()
    p = capi.default_params(**params)
    hip, ora = run_both(prob, p, capi.ORDERING_MIN_DEGREE)
    # same iteration sequence and errors along the way
    assert_same_trajectory(hip, ora, rtol=1e-8)
    np.testing.assert_allclose(hip["values"], ora["values"], rtol=0,
                               atol=1e-7 * max(1.0, np.abs(ora["values"]).max()))
