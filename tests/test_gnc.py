"""Graduated non-convexity on the CPU oracle (opt/gnc_optimizer.h:22-148) and the Barron factor
it drives (test/symforce_gnc_codegen_test.py:20-29, symforce/opt/noise_models.py:318-436).

The Barron loss is pinned by its closed-form members, written from their textbook definitions
rather than from the oracle's general expression: alpha = 1 (mu = 0, pseudo-Huber
sqrt(1 + e^2) - 1), alpha -> 0 (mu -> 1/2, Cauchy log(1 + e^2/2)) and alpha = -2 (mu = 3/4,
Geman-McClure 2 e^2 / (e^2 + 4)), each up to the epsilon terms of the reference's form; the
whitening r = sqrt(2 rho + eps) - sqrt(eps) is checked through the cost, and the dual-number
Jacobian by central differences; alpha = 0 itself (mu = 1/2) cancels in the reference's
form, so there the losses at alpha = +/-0.004 bracket the Cauchy loss. The GNC schedule is pinned by the reference's own "Test GNC"
known answers on its samples (tests/golden/gnc_problem.json, drawn as the reference draws them):
9 iteration records, |x| < 0.1, five times closer to 0 than the plain optimum at mu = 0, and
SUCCESS."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from problem_builders import gnc_problem, gnc_reference_samples

EPS = 10 * np.finfo(np.float64).eps


def barron_residual(x, y, mu, eps=EPS):
    out, _ = ow.linearize_factor(capi.FACTOR_BARRON_VECTOR5, [x, y, [mu], [eps]])
    r = out[0][:5]
    J = out[1][:25].reshape(5, 5, order="F")
    return r, J


def loss_from_residual(r, eps=EPS):
    # r = sqrt(2 rho + eps) - sqrt(eps)  <=>  rho = ((r + sqrt(eps))^2 - eps) / 2
    return ((r + np.sqrt(eps)) ** 2 - eps) / 2


@pytest.mark.parametrize("mu,closed_form,rtol", [
    (0.0, lambda e: np.sqrt(1 + e * e) - 1, 1e-12),
    (0.75, lambda e: 2 * e * e / (e * e + 4), 1e-12),
])
def test_barron_members_match_their_closed_forms(mu, closed_form, rtol):
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = rng.standard_normal(5) * 3
        y = rng.standard_normal(5)
        r, _ = barron_residual(x, y, mu)
        e = x - y
        np.testing.assert_allclose(loss_from_residual(r), closed_form(e), rtol=rtol, atol=1e-14)
        assert np.all(r >= 0)


def test_barron_brackets_cauchy_around_mu_one_half():
    # alpha = 0 itself is the family's removable singularity (the reference's form cancels to
    # ~1e-2 there in double precision): mu = 0.5 -/+ 1e-3 gives alpha = +/-0.004, whose losses
    # bracket the Cauchy loss log(1 + e^2/2) (the family increases with alpha) within 1%
    rng = np.random.default_rng(4)
    for _ in range(20):
        x, y = rng.standard_normal(5) * 3, rng.standard_normal(5)
        e = x - y
        lo = loss_from_residual(barron_residual(x, y, 0.501)[0])
        hi = loss_from_residual(barron_residual(x, y, 0.499)[0])
        cauchy = np.log1p(e * e / 2)
        assert np.all(lo <= cauchy * (1 + 1e-12)) and np.all(cauchy <= hi * (1 + 1e-12))
        np.testing.assert_allclose(lo, cauchy, rtol=1e-2)
        np.testing.assert_allclose(hi, cauchy, rtol=1e-2)


def test_barron_is_quadratic_near_zero_and_vanishes_at_zero():
    # every member is e^2 / 2 + O(e^4) (delta = 1): the GNC stages share the inlier basin
    for mu in (0.0, 0.33, 0.66, 0.99):
        r, _ = barron_residual(np.full(5, 1e-4), np.zeros(5), mu)
        np.testing.assert_allclose(loss_from_residual(r), 0.5e-8, rtol=1e-6)
        r0, _ = barron_residual(np.ones(5), np.ones(5), mu)
        np.testing.assert_array_equal(r0, 0.0)


@pytest.mark.parametrize("mu", [0.0, 0.33, 0.66, 0.99])
def test_barron_jacobian_matches_finite_differences(mu):
    rng = np.random.default_rng(int(mu * 100))
    x = rng.standard_normal(5) * 2
    y = rng.standard_normal(5)
    _, J = barron_residual(x, y, mu)
    h = 1e-6
    Jfd = np.zeros((5, 5))
    for k in range(5):
        d = np.zeros(5)
        d[k] = h
        Jfd[:, k] = (barron_residual(x + d, y, mu)[0] - barron_residual(x - d, y, mu)[0]) / (2 * h)
    np.testing.assert_allclose(J, Jfd, rtol=1e-6, atol=1e-8)
    assert np.count_nonzero(J - np.diag(np.diag(J))) == 0


def reference_gnc_params():
    g = capi.GncParams()
    g.mu_initial, g.mu_max, g.mu_step, g.gnc_update_min_reduction = 0.0, 0.99, 0.33, 1e-3
    return g


def test_oracle_gnc_reproduces_the_reference_known_answers():
    prob, u = gnc_problem(gnc_reference_samples())
    op = ow.OracleProblem(prob)
    res = op.optimize_gnc(reference_gnc_params(), u, params=capi.default_params(), epsilon=1e-12)
    x = res["values"][:5]
    plain = op.optimize(capi.default_params(), epsilon=1e-12)  # 'u' = 0 in the problem values
    x_plain = plain["values"][:5]
    assert len(res["iterations"]) == 9
    assert np.linalg.norm(x) < 0.1
    assert np.linalg.norm(x) * 5 < np.linalg.norm(x_plain)
    assert res["status"] == capi.STATUS_SUCCESS
    assert res["values"][-1] == pytest.approx(0.99)


@pytest.mark.parametrize("schedule", [(0.0, 0.0, 0.99), (0.4, 0.33, 0.4), (0.7, 0.1, 0.6)])
def test_oracle_gnc_without_a_schedule_is_one_plain_optimization(schedule):
    mu0, step, mu_max = schedule
    prob, u = gnc_problem(gnc_reference_samples())
    op = ow.OracleProblem(prob)
    g = capi.GncParams()
    g.mu_initial, g.mu_step, g.mu_max, g.gnc_update_min_reduction = mu0, step, mu_max, 1e-1
    res = op.optimize_gnc(g, u, params=capi.default_params(), epsilon=1e-12)
    vals = prob.values.copy()
    vals[u_storage(prob, u)] = mu0
    plain = op.optimize(capi.default_params(), epsilon=1e-12, values=vals)
    assert [it["new_error"] for it in res["iterations"]] == \
        [it["new_error"] for it in plain["iterations"]]
    np.testing.assert_array_equal(res["values"], plain["values"])


def u_storage(prob, u):
    sd = np.array([capi.STORAGE_DIM[int(t)] for t in prob.types])
    return int(sd[:u].sum())
