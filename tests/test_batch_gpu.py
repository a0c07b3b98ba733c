"""Batched persistent-kernel LM (many small robot-3D-localization problems in one launch; no
reference counterpart: the reference runs one Optimizer per thread, opt/optimizer.h:23).

Each problem's trajectory is held to the oracle's LM (the restated Iterate,
levenberg_marquardt_solver.tcc:138-343) on the same inputs, the reference robot-3D example's
known answer is reproduced inside a batch, and the FP32 form reaches the FP64 oracle's cost."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
import problem_builders as pb
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu


def best_error(r):
    return r["iterations"][r["best_index"]]["new_error"]


@pytest.mark.parametrize("lam", [capi.LAMBDA_STATIC, capi.LAMBDA_DYNAMIC])
def test_batch_matches_oracle_per_problem(lam):
    probs = capi.gen_localization_batch(24, seed=3)
    params = capi.default_params(lambda_update_type=lam, iterations=30)
    res = capi.HipBatch(probs, params).solve()
    for prob, got in zip(probs, res["problems"]):
        ora = ow.OracleProblem(prob).optimize(params)
        assert_same_trajectory(got, ora, rtol=1e-8)
        np.testing.assert_allclose(got["values"], ora["values"], rtol=0,
                                   atol=1e-8 * max(1.0, np.abs(ora["values"]).max()))
    assert (res["solve_ms"] > 0).all() and res["total_ms"] > 0


def test_reference_known_answer_inside_a_batch():
    """test/symforce_examples_robot_3d_localization_test.py:49-51, with other problems around."""
    prob, _ = pb.robot3d_problem()
    params = capi.default_params(initial_lambda=1e4, lambda_down_factor=0.5)
    batch = capi.gen_localization_batch(5, seed=9)[:2] + [prob] + capi.gen_localization_batch(3, seed=4)
    got = capi.HipBatch(batch, params).solve()["problems"][2]
    assert got["iterations"][0]["new_error"] == pytest.approx(463700.5576620833, abs=5e-7)
    assert best_error(got) < 140
    assert got["status"] == capi.STATUS_SUCCESS
    assert_same_trajectory(got, ow.OracleProblem(prob).optimize(params), rtol=1e-8)


def test_batch_agrees_with_single_problem_optimizer():
    probs = capi.gen_localization_batch(6, seed=12)
    params = capi.default_params(iterations=25)
    res = capi.HipBatch(probs, params).solve()["problems"]
    for prob, got in zip(probs, res):
        one = capi.HipOptimizer(prob, params).optimize()
        assert_same_trajectory(got, one, rtol=1e-9)


def test_fp32_form_reaches_fp64_cost():
    """FP32 Hessian blocks, FP64 residual, gradient and cost: under the default early exit every
    problem converges (no stall at the iteration limit), and run to convergence the FP32 form
    lands on the FP64 oracle's minimum, which the FP64 gradient fixes; only the steps on the way
    differ."""
    probs = capi.gen_localization_batch(32, seed=5)
    res = capi.HipBatch(probs, capi.default_params(iterations=50),
                        precision=capi.PRECISION_F32_JAC).solve()["problems"]
    assert all(got["status"] == capi.STATUS_SUCCESS for got in res)
    params = capi.default_params(iterations=40, early_exit_min_reduction=0.0)
    res = capi.HipBatch(probs, params, precision=capi.PRECISION_F32_JAC).solve()["problems"]
    for prob, got in zip(probs, res):
        ora = ow.OracleProblem(prob).optimize(params)
        assert best_error(got) == pytest.approx(best_error(ora), rel=1e-8)


def test_unsupported_problems_are_rejected():
    with pytest.raises(RuntimeError, match="matching and odometry"):
        capi.HipBatch([capi.gen_pose_smoothing(10, 42)])
    with pytest.raises(RuntimeError, match="1..64"):
        capi.HipBatch([capi.gen_localization(70, 10, 0.3, 1)])


@pytest.mark.parametrize("lam", [capi.LAMBDA_STATIC, capi.LAMBDA_DYNAMIC])
def test_full_batch_of_4096_problems_matches_oracle(lam):
    """The benchmark's batched workload at its full size: 4096 localization problems of 5-50
    poses and 5-40 landmarks in one launch, every problem's trajectory held to the oracle."""
    probs = capi.gen_localization_batch(4096, seed=42)
    params = capi.default_params(lambda_update_type=lam)
    res = capi.HipBatch(probs, params).solve()
    assert len(res["problems"]) == 4096
    sizes = [len(p.values) for p in probs]
    assert min(sizes) < 300 < max(sizes)  # the uneven sizes the launch is meant for
    for prob, got in zip(probs, res["problems"]):
        ora = ow.OracleProblem(prob).optimize(params)
        assert_same_trajectory(got, ora, rtol=1e-8)
        np.testing.assert_allclose(got["values"], ora["values"], rtol=0,
                                   atol=1e-8 * max(1.0, np.abs(ora["values"]).max()))


@pytest.mark.parametrize("precision", [capi.PRECISION_F64, capi.PRECISION_F32_JAC])
def test_every_launch_configuration_gives_the_same_results(precision):
    """The autotuner's configurations (occupancy one or two wavefronts per SIMD, problems as
    given or most poses first) change where and when problems run, not their arithmetic: every
    configuration's solve is bit-identical, and tuning keeps one of them."""
    probs = capi.gen_localization_batch(300, seed=8)
    b = capi.HipBatch(probs, capi.default_params(iterations=50), precision=precision)
    ref = b.solve()
    for c in range(4):
        b.set_config(c)
        r = b.solve()
        for a, x in zip(ref["problems"], r["problems"]):
            np.testing.assert_array_equal(a["values"], x["values"])
            assert (a["status"], a["best_index"]) == (x["status"], x["best_index"])
            assert [it["new_error"] for it in a["iterations"]] == \
                [it["new_error"] for it in x["iterations"]]
    chosen, ms = b.tune(reps=1)
    assert 0 <= chosen < 4 and np.all(ms > 0) and ms[chosen] == ms.min()
    r = b.solve()
    for a, x in zip(ref["problems"], r["problems"]):
        np.testing.assert_array_equal(a["values"], x["values"])
