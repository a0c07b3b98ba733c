"""LM branches beyond the default parameters, per iteration against the FP64 oracle:

* diagonal damping, with and without the running maximum of the diagonal (DampHessian,
  levenberg_marquardt_solver.tcc:22-54), with and without unit damping, on the sparse Cholesky
  path, the Schur path and the batched kernel;
* bold updates (the angle test of Iterate, levenberg_marquardt_solver.tcc:270-280) on the
  sparse Cholesky and Schur paths. The batched kernel rejects them (checked below).

Each case also asserts that the branch it is about was taken (a damping mode that changes the
trajectory; a non-zero update angle and an update accepted without reducing the error), so a
case cannot pass by never reaching the code it names."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu

DAMPING = {
    "diag": dict(use_diagonal_damping=1, use_unit_damping=0),
    "diag_unit": dict(use_diagonal_damping=1, use_unit_damping=1),
    "diag_keep_max": dict(use_diagonal_damping=1, keep_max_diagonal_damping=1,
                          use_unit_damping=0),
    "diag_keep_max_unit": dict(use_diagonal_damping=1, keep_max_diagonal_damping=1,
                               use_unit_damping=1, diagonal_damping_min=1e-3),
}

CHOLESKY_CASES = {
    "pose_smoothing": lambda: capi.gen_pose_smoothing(10, 42),
    "ba_circle": lambda: capi.gen_ba_circle(8, 150, 3),
    "pose_graph": lambda: capi.gen_pose_graph(2000, 150, 7),
}


def _errors(r):
    return np.array([it["new_error"] for it in r["iterations"]])


def _hip(prob, params, solver=capi.SOLVER_CHOLESKY):
    return capi.HipOptimizer(prob, params, capi.solver_options(
        linear_solver=solver, ordering=capi.ORDERING_MIN_DEGREE)).optimize()


def _oracle(prob, params, solver=ow.CHOLESKY):
    return ow.OracleProblem(prob).optimize(params, solver=solver,
                                           c_dim=int(prob.info.get("c_dim", 0)))


def _differs_from_simpler_damping(prob, params, got, solver=ow.CHOLESKY):
    """The oracle's trajectory under `params` differs from the one with the running maximum
    switched off (keep-max modes) or with plain unit damping (the other modes)."""
    if params.keep_max_diagonal_damping:
        simpler = capi.OptimizerParams.from_buffer_copy(params)
        simpler.keep_max_diagonal_damping = 0
    else:
        simpler = capi.default_params(iterations=params.iterations,
                                      lambda_update_type=params.lambda_update_type)
    a, b = _errors(got), _errors(_oracle(prob, simpler, solver))
    n = min(a.size, b.size)
    return a.size != b.size or not np.allclose(a[:n], b[:n], rtol=1e-6)


@pytest.mark.parametrize("mode", sorted(DAMPING))
@pytest.mark.parametrize("name", sorted(CHOLESKY_CASES))
def test_diagonal_damping_cholesky_path(name, mode):
    prob = CHOLESKY_CASES[name]()
    params = capi.default_params(iterations=30, **DAMPING[mode])
    hip, ora = _hip(prob, params), _oracle(prob, params)
    assert_same_trajectory(hip, ora, rtol=1e-8)
    np.testing.assert_allclose(hip["values"], ora["values"], rtol=0,
                               atol=1e-7 * max(1.0, np.abs(ora["values"]).max()))
    assert _differs_from_simpler_damping(prob, params, ora)


@pytest.mark.parametrize("mode", sorted(DAMPING))
@pytest.mark.parametrize("lam", [capi.LAMBDA_STATIC, capi.LAMBDA_DYNAMIC])
def test_diagonal_damping_schur_path(mode, lam):
    prob = capi.gen_ba_snavely(12, 300, 5)
    params = capi.default_params(iterations=30, lambda_update_type=lam, **DAMPING[mode])
    hip, ora = _hip(prob, params, capi.SOLVER_SCHUR), _oracle(prob, params, ow.SCHUR)
    assert_same_trajectory(hip, ora, rtol=1e-8)
    np.testing.assert_allclose(hip["values"], ora["values"], rtol=0,
                               atol=1e-7 * max(1.0, np.abs(ora["values"]).max()))
    assert _differs_from_simpler_damping(prob, params, ora, ow.SCHUR)


@pytest.mark.parametrize("mode", sorted(DAMPING))
def test_diagonal_damping_batched_kernel(mode):
    probs = capi.gen_localization_batch(16, seed=21)
    params = capi.default_params(iterations=30, **DAMPING[mode])
    res = capi.HipBatch(probs, params).solve()["problems"]
    changed = 0
    for prob, got in zip(probs, res):
        ora = ow.OracleProblem(prob).optimize(params)
        assert_same_trajectory(got, ora, rtol=1e-8)
        np.testing.assert_allclose(got["values"], ora["values"], rtol=0,
                                   atol=1e-8 * max(1.0, np.abs(ora["values"]).max()))
        changed += _differs_from_simpler_damping(prob, params, ora)
    assert changed > 0


def _bold_taken(ora):
    its = ora["iterations"]
    angles = sum(1 for it in its if it["update_angle_change"] != 0)
    bold = sum(1 for it in its if it["update_accepted"] and it["relative_reduction"] <= 0)
    return angles, bold


@pytest.mark.parametrize("name,gen,solver,lam", [
    ("pose_graph_static", lambda: capi.gen_pose_graph(2000, 150, 7), "cholesky",
     capi.LAMBDA_STATIC),
    ("pose_graph_dynamic", lambda: capi.gen_pose_graph(2000, 150, 7), "cholesky",
     capi.LAMBDA_DYNAMIC),
    ("ba_snavely_dynamic", lambda: capi.gen_ba_snavely(12, 300, 5), "schur",
     capi.LAMBDA_DYNAMIC),
])
def test_bold_updates(name, gen, solver, lam):
    prob = gen()
    params = capi.default_params(iterations=40, enable_bold_updates=1, lambda_update_type=lam)
    if solver == "cholesky":
        hip, ora = _hip(prob, params), _oracle(prob, params)
    else:
        hip, ora = _hip(prob, params, capi.SOLVER_SCHUR), _oracle(prob, params, ow.SCHUR)
    angles, bold = _bold_taken(ora)
    assert angles > 0 and bold > 0, (angles, bold)
    # The device orders the pose graph by minimum degree and the oracle naturally, so the errors
    # agree to rounding (~3e-11 relative here). DYNAMIC lambda scales by a cubic in the gain
    # ratio (e0 - e1) / (e0 - e_lin), which magnifies that relative error by e0 / |e0 - e1|
    # (1e3-1e4 on the late, small steps): lambda is held to 1e-6.
    assert_same_trajectory(hip, ora, rtol=1e-8, lambda_rtol=1e-6)
    for a, b in zip(hip["iterations"], ora["iterations"]):
        assert a["update_angle_change"] == pytest.approx(b["update_angle_change"], rel=1e-8,
                                                         abs=1e-12)


def test_bold_updates_rejected_by_batched_kernel():
    params = capi.default_params(enable_bold_updates=1)
    with pytest.raises(RuntimeError, match="bold"):
        capi.HipBatch(capi.gen_localization_batch(2, seed=1), params)
