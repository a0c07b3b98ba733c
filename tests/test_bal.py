"""Bundle Adjustment in the Large reader (the reference example's ReadProblem,
examples/bundle_adjustment_in_the_large/bundle_adjustment_in_the_large.cc:60-118).

No BAL data file ships with the reference (the example downloads them), so the reader is checked
on files written from seeded problems: the written text round-trips to the same Values, the
angle-axis -> Pose3 conversion agrees with Rodrigues' rotation formula, and the key layout and
factor arguments are the reference example's."""
import numpy as np
import pytest

from symforce_amd import _capi as capi

EPS = 10 * np.finfo(float).eps


def quat_to_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def rodrigues(r):
    th = np.linalg.norm(r)
    k = r / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def key_values(prob, letter):
    """{super: values} of every key with this letter."""
    width = {capi.TYPE_POSE3: 7, capi.TYPE_VECTOR3: 3, capi.TYPE_VECTOR2: 2, capi.TYPE_SCALAR: 1,
             capi.TYPE_MATRIX66: 36}
    out, off = {}, 0
    for k in range(len(prob.letters)):
        w = width.get(int(prob.types[k]), None)
        assert w is not None
        if chr(prob.letters[k]) == letter:
            out[int(prob.supers[k])] = prob.values[off:off + w]
        off += w
    return out


def write_small_bal(path):
    rng = np.random.default_rng(5)
    nc, npt = 3, 6
    obs = [(c, p) for p in range(npt) for c in range(nc) if (c + p) % 3 != 2]
    with open(path, "w") as f:
        f.write(f"{nc} {npt} {len(obs)}\n")
        for c, p in obs:
            f.write(f"{c} {p} {rng.normal() * 100:.17g} {rng.normal() * 100:.17g}\n")
        cams = []
        for c in range(nc):
            cam = np.concatenate([rng.normal(size=3) * 0.5, rng.normal(size=3), [500.0, 1e-3, -2e-4]])
            cams.append(cam)
            f.write(" ".join(f"{v:.17g}" for v in cam) + "\n")
        for p in range(npt):
            f.write(" ".join(f"{v:.17g}" for v in rng.normal(size=3)) + "\n")
    return cams, obs


def test_reader_layout_and_rotation(tmp_path):
    path = tmp_path / "small.txt"
    cams, obs = write_small_bal(path)
    prob = capi.load_bal(path)
    assert prob.info["n_cams"] == 3 and prob.info["n_points"] == 6 and prob.info["n_obs"] == len(obs)
    poses = key_values(prob, "c")
    for c, cam in enumerate(cams):
        q = poses[c][:4]
        assert np.linalg.norm(q) == pytest.approx(1.0, abs=1e-15)
        np.testing.assert_allclose(quat_to_matrix(q), rodrigues(cam[:3]), atol=1e-14)
        np.testing.assert_array_equal(poses[c][4:], cam[3:6])
        np.testing.assert_array_equal(key_values(prob, "i")[c], cam[6:])
    # one SnavelyReprojectionFactor per observation: (c_cam, i_cam, p_pt, P_obs, e)
    assert (prob.ftypes == capi.FACTOR_SNAVELY).all() and len(prob.ftypes) == len(obs)
    for o, (c, p) in enumerate(obs):
        a = prob.fargs[prob.fptr[o]:prob.fptr[o + 1]]
        got = [(chr(prob.letters[k]), int(prob.supers[k])) for k in a]
        assert got[:4] == [("c", c), ("i", c), ("p", p), ("P", o)]
        assert chr(prob.letters[a[4]]) == "e"
    assert key_values(prob, "e")[np.iinfo(np.int64).min][0] == EPS


def test_round_trip_through_bal_text(tmp_path):
    prob = capi.gen_ba_snavely(6, 60, 3)
    path = tmp_path / "gen.txt"
    capi.save_bal(prob, path)
    back = capi.load_bal(path, huber_k=2.0)
    a, b = key_values(prob, "c"), key_values(back, "c")
    for c in a:
        qa, qb = a[c][:4], b[c][:4]
        np.testing.assert_allclose(quat_to_matrix(qb), quat_to_matrix(qa), atol=1e-14)
        np.testing.assert_array_equal(a[c][4:], b[c][4:])
    for letter in "ip":
        va, vb = key_values(prob, letter), key_values(back, letter)
        assert va.keys() == vb.keys()
        for k in va:
            np.testing.assert_array_equal(va[k], vb[k])
    pa = [v for k, v in sorted(key_values(prob, "x").items())]
    pb = [v for k, v in sorted(key_values(back, "P").items())]
    np.testing.assert_array_equal(np.array(pa), np.array(pb))
    assert (back.ftypes == capi.FACTOR_SNAVELY_HUBER).all()


def test_malformed_file_is_an_error(tmp_path):
    path = tmp_path / "bad.txt"
    path.write_text("2 1 1\n5 0 1.0 2.0\n")
    with pytest.raises(RuntimeError):
        capi.load_bal(path)
    with pytest.raises(RuntimeError):
        capi.load_bal(tmp_path / "missing.txt")


@pytest.mark.gpu
def test_bal_problem_optimizes_like_the_oracle(tmp_path):
    """The reference example's setup: plain Snavely factors, no gauge prior, DYNAMIC lambda."""
    import oracle_wrap as ow
    from test_optimizer_gpu import assert_same_trajectory
    path = tmp_path / "ba.txt"
    capi.save_bal(capi.gen_ba_snavely(12, 300, 5), path)
    prob = capi.load_bal(path)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=20)
    hip = capi.HipOptimizer(prob, params,
                            capi.solver_options(linear_solver=capi.SOLVER_SCHUR)).optimize()
    ora = ow.OracleProblem(prob).optimize(params, solver=ow.SCHUR, c_dim=int(prob.info["c_dim"]))
    # the free 7-dof similarity gauge leaves H damped only by lambda: ~1e-11 cost differences,
    # which the DYNAMIC lambda update amplifies through the gain ratio
    assert_same_trajectory(hip, ora, rtol=1e-7, lambda_rtol=1e-6)
