"""Pins the oracle's robust losses and reprojection bodies independently of their C++ formulas.

The oracle whitens a residual vector x by its norm (symforce/opt/noise_models.py:68-90):
w(x) = f(n)/n * x with n = sqrt(|x|^2 + eps) and f(n) = sqrt(2 rho(n)), so the factor's cost
1/2 |w|^2 must equal rho(n) * |x|^2 / n^2. Each check below recomputes the unwhitened residual
in numpy from the textbook projection formulas and compares the oracle's whitened residual with
a loss written from its definition, not from the oracle's expression:

* Huber (k): the classical piecewise loss x^2/2 for |x| <= k, k|x| - k^2/2 beyond, with the
  identity below the switch and continuity of value and slope at |x| = k.
* Cauchy (c): SymForce's Barron loss (noise_models.py:318-428) at alpha = -1e-4 and +1e-4 with
  delta = c / sqrt(2) brackets the oracle's value (the family is increasing in alpha and its
  alpha -> 0 member is the Cauchy loss delta^2 log(1 + x^2 / (2 delta^2))).
* The reprojection bodies (BAL Snavely camera, bundle_adjustment_in_the_large.py:18-63; linear
  pinhole, gen/cpp/sym/linear_camera_cal.cc:49-72) vanish at noiseless pixels and equal the
  numpy projection elsewhere.
* The dual-number Jacobians of the whitened factors equal central finite differences through
  the Pose3 retraction (gen/cpp/sym/ops/pose3/lie_group_ops.cc:69-102).
"""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow

EPS = 10 * np.finfo(np.float64).eps


def quat_to_rot(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz])


def pose_retract(p, d):
    """Pose3 retraction of the reference: rotation right-multiplied by exp(d[:3]), t += d[3:]."""
    w = np.asarray(d[:3], float)
    th = np.sqrt(EPS ** 2 + w @ w)
    dq = np.concatenate([np.sin(th / 2) / th * w, [np.cos(th / 2)]])
    return np.concatenate([quat_mul(p[:4], dq), p[4:] + d[3:]])


def random_pose(rng, t_scale=1.0):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    return np.concatenate([q, t_scale * rng.standard_normal(3)])


def snavely_pixel(cam_T_world, intr, point):
    pc = quat_to_rot(cam_T_world[:4]) @ point + cam_T_world[4:]
    p = -pc[:2] / pc[2]
    r2 = p @ p
    return intr[0] * (1 + intr[1] * r2 + intr[2] * r2 * r2) * p


def linear_pixel(world_T_cam, point, K):
    pc = quat_to_rot(world_T_cam[:4]).T @ (point - world_T_cam[4:])
    z = max(pc[2], EPS)
    return np.array([K[0] * pc[0] / z + K[2], K[1] * pc[1] / z + K[3]])


def huber_rho(n, k):
    return np.where(n <= k, 0.5 * n * n, k * n - 0.5 * k * k)


def barron_rho(x, alpha, delta, alpha_eps=1e-12):
    """noise_models.py:386-399 (scalar_information = 1); (1 + u)^(d/2) - 1 is evaluated as
    expm1(d/2 log1p(u)), the same function without the cancellation at small d."""
    b = abs(alpha - 2) + alpha_eps
    d = alpha + (1.0 if alpha >= 0 else -1.0) * alpha_eps
    return delta ** 2 * (b / d) * np.expm1(d / 2 * np.log1p(x * x / delta ** 2 / b))


def factor_residual(ftype, args):
    out, _ = ow.linearize_factor(ftype, args)
    return out


def snavely_case(rng):
    """A camera looking down -z at a point in front of it (BAL sign: X_c,z < 0)."""
    cam = random_pose(rng)
    R = quat_to_rot(cam[:4])
    pc = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), -rng.uniform(3, 8)])
    point = R.T @ (pc - cam[4:])
    intr = np.array([rng.normal(500, 20), rng.normal(0, 1e-3), rng.normal(0, 1e-3)])
    return cam, intr, point


def test_snavely_body_matches_numpy_projection():
    rng = np.random.default_rng(3)
    for _ in range(20):
        cam, intr, point = snavely_case(rng)
        clean = snavely_pixel(cam, intr, point)
        r0 = factor_residual(capi.FACTOR_SNAVELY, [cam, intr, point, clean, [EPS]])[0][:2]
        assert np.abs(r0).max() < 1e-9
        pix = clean + rng.normal(0, 3, 2)
        r = factor_residual(capi.FACTOR_SNAVELY, [cam, intr, point, pix, [EPS]])[0][:2]
        np.testing.assert_allclose(r, clean - pix, rtol=1e-10, atol=1e-9)


def test_linear_camera_body_matches_numpy_projection():
    rng = np.random.default_rng(4)
    K = np.array([500.0, 500.0, 320.0, 240.0])
    for _ in range(20):
        cam = random_pose(rng, 5.0)
        pc = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(3, 8)])
        point = quat_to_rot(cam[:4]) @ pc + cam[4:]
        clean = linear_pixel(cam, point, K)
        # the Cauchy whitening scales a residual near zero by ~1: the body vanishes through it
        r0 = factor_residual(capi.FACTOR_REPROJ_CAUCHY, [cam, point, clean, K, [1.0], [EPS]])
        assert np.abs(r0[0][:2]).max() < 1e-9


@pytest.mark.parametrize("k", [0.5, 2.0, 7.0])
def test_huber_whitening_is_the_huber_loss(k):
    rng = np.random.default_rng(5)
    cam, intr, point = snavely_case(rng)
    clean = snavely_pixel(cam, intr, point)
    norms = np.concatenate([np.geomspace(1e-3, 50.0, 40) * k / 2.0, k * (1 + np.array(
        [-1e-9, -1e-12, 0.0, 1e-12, 1e-9, 1e-6]))])
    for n_target in norms:
        direction = rng.standard_normal(2)
        direction /= np.linalg.norm(direction)
        pix = clean - n_target * direction
        x = factor_residual(capi.FACTOR_SNAVELY, [cam, intr, point, pix, [EPS]])[0][:2]
        w = factor_residual(capi.FACTOR_SNAVELY_HUBER, [cam, intr, point, pix, [k], [EPS]])[0][:2]
        n = np.sqrt(x @ x + EPS)
        want = huber_rho(n, k) * (x @ x) / (n * n)
        assert 0.5 * (w @ w) == pytest.approx(want, rel=1e-12, abs=1e-300)
        # whitening only rescales: w is a non-negative multiple of x
        assert abs(w[0] * x[1] - w[1] * x[0]) <= 1e-12 * (np.abs(w).max() * np.abs(x).max())
        assert w @ x >= 0
        if n <= k:
            np.testing.assert_array_equal(w, x)
    # value and slope of the loss are continuous at the switch
    lo, hi = huber_rho(k * (1 - 1e-9), k), huber_rho(k * (1 + 1e-9), k)
    assert hi - lo == pytest.approx(2e-9 * k * k, rel=1e-5)


@pytest.mark.parametrize("c", [0.25, 1.0, 4.0])
def test_cauchy_whitening_is_the_barron_alpha_zero_limit(c):
    rng = np.random.default_rng(6)
    K = np.array([500.0, 480.0, 320.0, 240.0])
    cam = random_pose(rng, 5.0)
    pc = np.array([0.2, -0.3, 5.0])
    point = quat_to_rot(cam[:4]) @ pc + cam[4:]
    clean = linear_pixel(cam, point, K)
    delta = c / np.sqrt(2.0)
    for n_target in np.geomspace(1e-3, 1e3, 37) * c:
        direction = rng.standard_normal(2)
        direction /= np.linalg.norm(direction)
        pix = clean - n_target * direction
        x = linear_pixel(cam, point, K) - pix
        w = factor_residual(capi.FACTOR_REPROJ_CAUCHY, [cam, point, pix, K, [c], [EPS]])[0][:2]
        n = np.sqrt(x @ x + EPS)
        cost = 0.5 * (w @ w) * (n * n) / (x @ x)  # rho(n)
        # x is a difference of pixels of size |clean|: both sides round it at a few ulp of |clean|
        rel = 1e-12 + 2e-15 * np.abs(clean).max() / n
        lo, hi = barron_rho(n, -1e-4, delta), barron_rho(n, 1e-4, delta)
        assert lo <= cost * (1 + rel) and cost <= hi * (1 + rel), (n, lo, cost, hi)
        assert cost == pytest.approx(0.5 * (lo + hi), rel=1e-7 + rel)
        assert abs(w[0] * x[1] - w[1] * x[0]) <= rel * (np.abs(w).max() * np.abs(x).max())


def test_cauchy_between_factor_whitens_the_between_residual():
    rng = np.random.default_rng(7)
    S = np.diag(rng.uniform(0.5, 3.0, 6))
    for c in (0.3, 2.0):
        for _ in range(10):
            a, b, m = random_pose(rng), random_pose(rng), random_pose(rng)
            x = factor_residual(capi.FACTOR_BETWEEN_POSE3, [a, b, m, S, [EPS]])[0][:6]
            w = factor_residual(capi.FACTOR_BETWEEN_POSE3_CAUCHY, [a, b, m, S, [c], [EPS]])[0][:6]
            n = np.sqrt(x @ x + EPS)
            want = 0.5 * c * c * np.log1p(n * n / (c * c)) * (x @ x) / (n * n)
            assert 0.5 * (w @ w) == pytest.approx(want, rel=1e-12)
            assert barron_rho(n, -1e-4, c / np.sqrt(2)) <= want * (x @ x + EPS) / (x @ x) * (
                1 + 1e-12) <= barron_rho(n, 1e-4, c / np.sqrt(2)) * (1 + 1e-12)


def _fd_jacobian(ftype, args, lin_args, tdims, h=1e-6):
    cols = []
    for ai, td in zip(lin_args, tdims):
        for j in range(td):
            d = np.zeros(td)
            d[j] = h
            plus, minus = [list(map(np.array, args)) for _ in range(2)]
            if td == 6 and len(args[ai]) == 7:
                plus[ai], minus[ai] = pose_retract(args[ai], d), pose_retract(args[ai], -d)
            else:
                plus[ai], minus[ai] = args[ai] + d, args[ai] - d
            rp = factor_residual(ftype, plus)[0]
            rm = factor_residual(ftype, minus)[0]
            cols.append((rp - rm) / (2 * h))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("which", ["huber", "cauchy_reproj", "cauchy_between"])
def test_whitened_jacobians_match_finite_differences(which):
    rng = np.random.default_rng(8)
    if which == "huber":
        cam, intr, point = snavely_case(rng)
        pix = snavely_pixel(cam, intr, point) + np.array([6.0, -4.0])  # beyond k = 2
        ftype, args, lin, tds, rd = (capi.FACTOR_SNAVELY_HUBER,
                                     [cam, intr, point, pix, np.array([2.0]), np.array([EPS])],
                                     [0, 1, 2], [6, 3, 3], 2)
    elif which == "cauchy_reproj":
        K = np.array([500.0, 500.0, 320.0, 240.0])
        cam = random_pose(rng, 5.0)
        point = quat_to_rot(cam[:4]) @ np.array([0.3, 0.1, 4.0]) + cam[4:]
        pix = linear_pixel(cam, point, K) + np.array([1.5, 0.7])
        ftype, args, lin, tds, rd = (capi.FACTOR_REPROJ_CAUCHY,
                                     [cam, point, pix, K, np.array([1.0]), np.array([EPS])],
                                     [0, 1], [6, 3], 2)
    else:
        S = np.diag(rng.uniform(0.5, 3.0, 6))
        ftype, args, lin, tds, rd = (capi.FACTOR_BETWEEN_POSE3_CAUCHY,
                                     [random_pose(rng), random_pose(rng), random_pose(rng), S,
                                      np.array([0.5]), np.array([EPS])], [0, 1], [6, 6], 6)
    out = factor_residual(ftype, args)
    td = sum(tds)
    J = out[1][: rd * td].reshape(td, rd).T
    J_fd = _fd_jacobian(ftype, args, lin, tds)[:rd]
    scale = np.abs(J).max()
    np.testing.assert_allclose(J, J_fd, rtol=0, atol=2e-6 * scale)
