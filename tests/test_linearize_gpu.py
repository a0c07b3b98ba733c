"""HIP linearization (residual, hessian_lower values in reference CSC order, rhs) against the FP64
oracle on the same seeded inputs. Tolerance: 1e-9 relative to each array's max magnitude, plus
1e-9 element-wise relative (FP64 mode)."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
import problem_builders as pb

from test_structure import CASES

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def assert_close(a, b, what, mag=None, atol=0.0):
    """|a - b| <= RTOL * (|b| + 1e-3 * mag) + atol: relative to each entry, with a floor set by
    the magnitude of the terms summed into it (cancellation in exact arithmetic leaves rounding
    noise of that size, not of the size of the result)."""
    mag = np.full(b.shape, float(np.max(np.abs(b), initial=0.0))) if mag is None else mag
    err = np.abs(a - b)
    bad = err > RTOL * (np.abs(b) + 1e-3 * mag) + atol
    assert not bad.any(), f"{what}: {bad.sum()} entries outside tolerance, max err {err.max()}"


def assert_linearization_close(got, want, colptr, rowind):
    n = colptr.size - 1
    diag = want["hessian_values"][colptr[:-1]]
    sd = np.sqrt(np.maximum(diag, 0.0))
    rnorm = float(np.linalg.norm(want["residual"]))
    # residuals that vanish in exact arithmetic (odometry-consistent chains) carry rounding noise
    # of eps * |sqrt_info| * |position| ~ 1e-12 for positions up to 1e3 m: floor of 1e-10
    assert_close(got["residual"], want["residual"], "residual", atol=1e-10)
    # |rhs_i| <= ||J_i|| ||r|| = sqrt(H_ii) ||r||: the size of the terms summed into rhs_i
    assert_close(got["rhs"], want["rhs"], "rhs", mag=sd * rnorm)
    cols = np.repeat(np.arange(n), np.diff(colptr))
    assert_close(got["hessian_values"], want["hessian_values"], "hessian",
                 mag=sd[rowind] * sd[cols])


def perturbed(prob, seed, scale=0.02):
    """Values moved off the generator's point so cancellations do not hide errors."""
    rng = np.random.default_rng(seed)
    v = prob.values.copy()
    return v + scale * rng.standard_normal(v.size) * (np.abs(v) > 0)


@pytest.mark.parametrize("name", sorted(CASES))
def test_linearize_matches_oracle(name):
    prob = CASES[name]()
    hp = capi.HipProblem(prob)
    op = ow.OracleProblem(prob)
    for values in (prob.values, perturbed(prob, 1)):
        # re-normalize quaternions after perturbation
        v = values.copy()
        for k, t in enumerate(prob.types):
            if t in (capi.TYPE_POSE3, capi.TYPE_ROT3):
                o = prob.storage_offsets[k]
                v[o:o + 4] /= np.linalg.norm(v[o:o + 4])
        got = hp.linearize(v)
        want = op.linearize(v)
        assert_linearization_close(got, want, *op.hessian_structure())


def test_linearize_snavely_mid_size():
    prob = capi.gen_ba_snavely(60, 3000, 11)
    hp = capi.HipProblem(prob)
    op = ow.OracleProblem(prob)
    assert_linearization_close(hp.linearize(), op.linearize(), *op.hessian_structure())
