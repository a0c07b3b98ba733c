"""Graduated non-convexity on the device (GncOptimizer, opt/gnc_optimizer.h:22-148) against the
oracle's restatement and the reference's "Test GNC" known answers (test/symforce_gnc_test.cc):

* sym::GncOptimizer<sym::Optimizerd> in C++ on the reference problem, its samples drawn as the
  reference draws them: 9 records, |x| < 0.1, five times closer to 0 than the plain optimum,
  SUCCESS; the same samples as the committed fixture; the oracle's trajectory;
* the ctypes entry point on larger problems (many outliers, STATIC and DYNAMIC lambda, a
  schedule that runs out of iterations) against the oracle, record by record;
* schedules that do not step mu (mu_step <= 0, mu_initial >= mu_max) equal one plain
  optimization; a mu key that is optimized or not a Scalar is rejected."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from problem_builders import gnc_problem, gnc_reference_samples
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu
BUILD = Path(__file__).resolve().parent / "cpp" / "_build"


def gnc_params(mu0=0.0, mu_max=0.99, step=0.33, min_red=1e-3):
    g = capi.GncParams()
    g.mu_initial, g.mu_max, g.mu_step, g.gnc_update_min_reduction = mu0, mu_max, step, min_red
    return g


def run_both(prob, u, g, params, num_iterations=-1):
    hip = capi.HipOptimizer(prob, params, epsilon=1e-12).optimize_gnc(
        g, u, num_iterations=num_iterations)
    ora = ow.OracleProblem(prob).optimize_gnc(g, u, params=params, epsilon=1e-12,
                                              num_iterations=num_iterations)
    for it in ora["iterations"] + hip["iterations"]:
        it.pop("relative_reduction", None)
        it.pop("update_angle_change", None)
    return hip, ora


def test_cpp_gnc_optimizer_reference_known_answers():
    out = subprocess.run([str(BUILD / "template_api"), "gnc"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert len(r["iterations"]) == 9
    assert np.linalg.norm(r["x_gnc"]) < 0.1
    assert np.linalg.norm(r["x_gnc"]) * 5 < np.linalg.norm(r["x_regular"])
    assert r["status"] == capi.STATUS_SUCCESS
    assert r["mu_final"] == pytest.approx(0.99)
    ys = gnc_reference_samples()
    np.testing.assert_array_equal(np.array(r["ys"]).reshape(-1, 5), ys)
    prob, u = gnc_problem(ys)
    ora = ow.OracleProblem(prob).optimize_gnc(gnc_params(), u, params=capi.default_params(),
                                              epsilon=1e-12)
    hip = dict(status=r["status"], failure_reason=0, best_index=r["best_index"],
               iterations=r["iterations"])
    for it in ora["iterations"]:
        it.pop("relative_reduction")
        it.pop("update_angle_change")
    assert_same_trajectory(hip, ora, rtol=1e-10)
    np.testing.assert_allclose(r["x_gnc"], ora["values"][:5], rtol=0, atol=1e-10)


def outlier_problem(n, frac, seed, dim_scale=10.0):
    rng = np.random.default_rng(seed)
    ys = 0.1 * rng.standard_normal((n, 5))
    k = int(frac * n)
    ys[:k] += dim_scale * (1 + rng.random((k, 5)))
    return gnc_problem(ys, x0=rng.standard_normal(5) * 2)


@pytest.mark.parametrize("lam", ["static", "dynamic"])
@pytest.mark.parametrize("case", [(300, 0.2, 5), (1000, 0.35, 6)])
def test_gnc_matches_oracle(lam, case):
    prob, u = outlier_problem(*case)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC if lam == "dynamic"
                                 else capi.LAMBDA_STATIC, iterations=100)
    hip, ora = run_both(prob, u, gnc_params(step=0.2, min_red=1e-2), params)
    assert len(hip["iterations"]) > 5
    assert_same_trajectory(hip, ora, rtol=1e-9, lambda_rtol=1e-6)
    np.testing.assert_allclose(hip["values"], ora["values"], rtol=0, atol=1e-9)
    # graduated: the final mu is the maximum, and the robust solution ignores the outliers
    assert hip["values"][prob.values.size - 1] == pytest.approx(0.99)
    assert np.linalg.norm(hip["values"][:5]) < 0.1


def test_gnc_budget_runs_across_stages():
    prob, u = outlier_problem(300, 0.2, 7)
    params = capi.default_params(iterations=100)
    hip, ora = run_both(prob, u, gnc_params(step=0.05, min_red=1e-6), params, num_iterations=12)
    # the 12 iterations (13 records) are shared by the stages: several ran, mu never got to
    # its maximum
    assert len(hip["iterations"]) == 12
    assert 0.05 <= hip["values"][-1] < 0.99
    assert_same_trajectory(hip, ora, rtol=1e-9, lambda_rtol=1e-6)
    np.testing.assert_allclose(hip["values"], ora["values"], rtol=0, atol=1e-9)


# (mu = 1/2 itself is alpha = 0, where the reference's form cancels to a few ulps)
@pytest.mark.parametrize("schedule", [(0.0, 0.0, 0.99), (0.4, 0.33, 0.4), (0.7, 0.1, 0.6)])
def test_gnc_without_a_schedule_is_one_plain_optimization(schedule):
    mu0, step, mu_max = schedule
    prob, u = gnc_problem(gnc_reference_samples())
    g = gnc_params(mu0, mu_max, step, 1e-1)
    hip = capi.HipOptimizer(prob, capi.default_params(), epsilon=1e-12).optimize_gnc(g, u)
    vals = prob.values.copy()
    vals[-1] = mu0
    plain = capi.HipOptimizer(prob, capi.default_params(), epsilon=1e-12).optimize(vals)
    assert [it["new_error"] for it in hip["iterations"]] == \
        [it["new_error"] for it in plain["iterations"]]
    np.testing.assert_array_equal(hip["values"], plain["values"])


def test_gnc_rejects_a_bad_mu_key():
    prob, u = gnc_problem(gnc_reference_samples())
    opt = capi.HipOptimizer(prob, capi.default_params(), epsilon=1e-12)
    with pytest.raises(RuntimeError, match="Scalar"):
        opt.optimize_gnc(gnc_params(), 0)  # 'x' is a Vector5
    with pytest.raises(RuntimeError, match="range"):
        opt.optimize_gnc(gnc_params(), prob.n_keys)
