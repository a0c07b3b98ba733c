"""The C++ host surface (include/sym/symforce_amd.h) used the way the reference's own tests use
the reference API, run on the device and checked against the CPU oracle or the mathematics:

* sym::Optimizerd on "Test pose smoothing" (test/symforce_optimizer_test.cc:142-176);
* sym::LevenbergMarquardtSolver<Scalar> with a caller LinearizeFunc, double and float:
  "Converges for a linear problem in one iteration" (test/symforce_levenberg_marquardt_test.cc);
* sym::Optimizer<double, LevenbergMarquardtSolver<double, hip::SparseSchurSolver<>>> through the
  virtual Optimize overload, held to the oracle's SparseSchurSolver trajectory;
* debug_stats (levenberg_marquardt_solver.tcc:115-122, :205-212): per-iteration update, Values
  and residual against the oracle's, and the factor's ordering and L pattern against the oracle's
  up-looking LDL^T for the same permutation (bit-exact pattern);
* Linearize, ComputeFullCovariance, ComputeAllCovariances and the Schur-complement
  ComputeCovariances (opt/internal/covariance_utils.h:151-170) against numpy inverses;
* sym::Optimizerf on the pose-smoothing known answer."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu
BUILD = Path(__file__).resolve().parent / "cpp" / "_build"


def run(binary, *args):
    exe = BUILD / binary
    assert exe.exists(), "build with python -m symforce_amd.build"
    out = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_cpp_pose_smoothing_known_answer():
    r = run("pose_smoothing")
    assert r["iteration"] == 12
    assert r["lambda"] == pytest.approx(0.0039, rel=1e-1)
    assert r["error"] == pytest.approx(7.801, rel=1e-3)
    assert r["status"] == 1


@pytest.mark.parametrize("scalar", ["double", "float"])
def test_lm_solver_converges_on_a_linear_problem_in_one_iteration(scalar):
    r = run("template_api", "lm_linear", scalar)
    assert r["error_init"] > 10000.0
    assert r["error_final"] < 1e-8
    assert r["best_norm"] < 1e-4
    assert r["records"] == 2  # the iteration -1 record and the one iteration
    assert r["ordering_size"] == 6 and r["L_shape"] == [6, 6]  # debug_stats on the solver


def test_optimizer_template_with_schur_solver_matches_oracle():
    r = run("template_api", "ba_schur")
    prob = capi.gen_ba_snavely(12, 300, 5)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=30)
    ora = ow.OracleProblem(prob).optimize(params, solver=ow.SCHUR,
                                          c_dim=int(prob.info["c_dim"]))
    hip = dict(status=r["status"], failure_reason=r["failure_reason"],
               best_index=r["best_index"], iterations=r["iterations"])
    for it in ora["iterations"]:
        it.pop("relative_reduction")
        it.pop("update_angle_change")
    assert_same_trajectory(hip, ora, rtol=1e-8)
    np.testing.assert_allclose(r["values"], ora["values"], rtol=0,
                               atol=1e-7 * max(1.0, np.abs(ora["values"]).max()))
    assert r["num_keys"] == 12 * 2 + 300


def test_debug_stats_match_the_oracle():
    r = run("template_api", "debug_stats")
    prob = capi.gen_pose_smoothing(10, 42)
    params = capi.default_params(iterations=50, early_exit_min_reduction=1e-4)
    op = ow.OracleProblem(prob)
    ora = op.optimize(params, record_updates=True)
    its = r["iterations"]
    assert len(its) == len(ora["iterations"])
    np.testing.assert_array_equal(its[0]["values"], r["start"])
    assert its[0]["update"] == []
    for k, it in enumerate(its[1:]):
        want = ora["updates"][k]
        np.testing.assert_allclose(it["update"], want, rtol=0,
                                   atol=1e-9 * max(1.0, np.abs(want).max()))
    for it in its:
        # the residual recorded is the one of the Values recorded
        want = op.linearize(np.array(it["values"]))["residual"]
        np.testing.assert_allclose(it["residual"], want, rtol=0,
                                   atol=1e-9 * max(1.0, np.abs(want).max()))
        assert 0.5 * np.dot(it["residual"], it["residual"]) == pytest.approx(it["new_error"],
                                                                              rel=1e-12)
    # the factor pattern equals the reference's up-looking LDL^T pattern for this permutation
    n = len(r["ordering"])
    perm = np.array(r["ordering"])
    assert sorted(perm) == list(range(n))
    inv = np.empty(n, np.int32)
    inv[perm] = np.arange(n)
    cp, ri = op.hessian_structure()
    chol = ow.OracleCholesky(n, cp, ri, ordering=inv)
    chol.factorize(op.linearize()["hessian_values"])
    L = chol.factor()
    np.testing.assert_array_equal(r["L_colptr"], L["colptr"])
    np.testing.assert_array_equal(r["L_rowind"], L["rowind"])
    assert r["L_shape"] == [n, n]
    best = ora["values"]
    np.testing.assert_allclose(r["best_residual"], op.linearize(best)["residual"], rtol=0,
                               atol=1e-9)


def test_covariances_are_the_hessian_inverse_and_its_schur_complement():
    r = run("template_api", "covariance")
    n = r["n"]
    H = ow.csc_to_dense(n, np.array(r["colptr"]), np.array(r["rowind"]), np.array(r["values"]))
    H = H + np.tril(H, -1).T
    full = np.array(r["full"]).reshape(n, n).T
    want = np.linalg.inv(H)
    np.testing.assert_allclose(full, want, rtol=0, atol=1e-8 * np.abs(want).max())
    assert r["all_vs_full"] == 0.0
    assert r["prefix_info"] == 0
    # leading 5 keys (Pose3: 30 tangent dims): inverse of their Schur complement
    nb = 30
    B, E, C = H[:nb, :nb], H[:nb, nb:], H[nb:, nb:]
    S_inv = np.linalg.inv(B - E @ np.linalg.solve(C, E.T))
    for i, blk in enumerate(r["prefix_blocks"]):
        got = np.array(blk).reshape(6, 6).T
        np.testing.assert_allclose(got, S_inv[6 * i:6 * i + 6, 6 * i:6 * i + 6], rtol=0,
                                   atol=1e-8 * np.abs(S_inv).max())


def test_optimizerf_pose_smoothing():
    r = run("template_api", "optimizerf")
    assert r["status"] == 1
    assert r["iteration"] == 12
    assert r["error"] == pytest.approx(7.801, rel=1e-3)
