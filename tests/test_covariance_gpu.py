"""Covariances from the device LDL^T: Optimizer::ComputeFullCovariance / ComputeAllCovariances
(opt/optimizer.tcc:113-121, 202-206; LevenbergMarquardtSolver::ComputeCovariance,
opt/levenberg_marquardt_solver.tcc:345-356) on the best linearization of an optimized problem.

The reference has no covariance test with a known answer for these calls, so the mathematics
pins them: the full covariance is the inverse of the symmetric Hessian (H C = I, C symmetric),
and the per-key blocks are its diagonal blocks at the state offsets (tangent dims 6 for Pose3)."""
import numpy as np
import pytest

from symforce_amd import _capi as capi

pytestmark = pytest.mark.gpu


def dense_lower_to_full(cp, ri, vals, n):
    h = np.zeros((n, n))
    for j in range(n):
        for q in range(cp[j], cp[j + 1]):
            h[ri[q], j] = vals[q]
    return h + np.tril(h, -1).T


@pytest.fixture(scope="module")
def optimized():
    prob = capi.gen_pose_smoothing(10, 42)
    opt = capi.HipOptimizer(prob, capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC))
    opt.optimize()
    hp = capi.HipProblem(prob)
    lin = opt.best_linearization(hp.M, hp.nnz, hp.N)
    cp, ri = hp.hessian_structure()
    return hp, lin["hessian_values"], dense_lower_to_full(cp, ri, lin["hessian_values"], hp.N)


def test_full_covariance_is_the_hessian_inverse(optimized):
    hp, vals, h = optimized
    cov = capi.compute_full_covariance(hp, vals)
    assert cov.shape == (hp.N, hp.N)
    np.testing.assert_allclose(cov, cov.T, rtol=0, atol=1e-9 * np.abs(cov).max())
    np.testing.assert_allclose(h @ cov, np.eye(hp.N), rtol=0, atol=1e-9 * np.linalg.cond(h))
    np.testing.assert_allclose(cov, np.linalg.inv(h), rtol=1e-7, atol=1e-9 * np.abs(cov).max())


def test_all_covariances_are_the_diagonal_blocks(optimized):
    hp, vals, h = optimized
    blocks = capi.compute_all_covariances(hp, vals)
    keys, offsets = hp.state_keys()
    assert sorted(blocks) == sorted(int(k) for k in keys)
    full = np.linalg.inv(h)
    for k, o in zip(keys, offsets):
        b = blocks[int(k)]
        assert b.shape == (6, 6)  # Pose3 tangent
        np.testing.assert_allclose(b, full[o:o + 6, o:o + 6], rtol=1e-7,
                                   atol=1e-9 * np.abs(full).max())
        assert np.all(np.linalg.eigvalsh(b) > 0)
