"""Mixed-precision supernodal LDL^T, the pose-graph precision form (SYMHIP_PRECISION_F32_UPDATES):
the trailing supernode updates run on the FP32 matrix cores, the diagonal blocks, panel solves,
D and the triangular solves stay FP64, and every optimizer solve is refined against the FP64
damped Hessian by conjugate gradients preconditioned by the mixed factor (FP64 products,
residuals and scalars) to a relative residual of 1e-13 or until it stagnates at the system's
FP64 floor (accepted below 1e-10; the FP64 factorization itself reaches ~1e-11 on the
ill-conditioned late LM systems of these pose graphs); a solve that misses it refactors with
FP64 updates. No reference counterpart: the reference factorizes in the optimizer's Scalar
(sparse_cholesky_solver.tcc:108-219).

* The raw mixed factor (standalone handle, no refinement) is FP32-accurate: its solution is
  farther from the dense FP64 solution than the FP64 factor's by orders of magnitude (the FP32
  updates run) yet within FP32 rounding times the conditioning, and refinement steps on the host
  contract its error geometrically to the FP64 factor's accuracy.
* With refinement the optimizer follows the oracle's LM trajectory to the FP64 path's tolerance
  on pose-graph, circle-BA and localization problems and on the 100 000-node pose graph, and
  every solve's final relative residual is recorded and small."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from test_optimizer_gpu import assert_same_trajectory
from test_configs_gpu import nd_permutation

pytestmark = pytest.mark.gpu


def assert_same_trajectory_up_to_ties(hip, ora, rtol=1e-8, lambda_rtol=1e-8):
    """The oracle's trajectory to rtol; the best index may differ only where the oracle's own
    costs at the two indices tie to 1e-12 (records at the converged floor, which FP64 rounding
    orders either way)."""
    hb, ob = hip["best_index"], ora["best_index"]
    if hb != ob:
        e = [it["new_error"] for it in ora["iterations"]]
        assert e[hb] == pytest.approx(e[ob], rel=1e-12), (hb, ob, e[hb], e[ob])
        hip = dict(hip, best_index=ob)
    assert_same_trajectory(hip, ora, rtol=rtol, lambda_rtol=lambda_rtol)


def mixed_options(ordering=capi.ORDERING_NESTED_DISSECTION):
    return capi.solver_options(linear_solver=capi.SOLVER_CHOLESKY, ordering=ordering,
                               precision=capi.PRECISION_F32_UPDATES)


def test_raw_mixed_factor_is_fp32_accurate_and_refines_to_fp64():
    prob = capi.gen_pose_graph(400, 60, 3)
    op = ow.OracleProblem(prob)
    cp, ri = op.hessian_structure()
    n = len(cp) - 1
    vals = op.linearize()["hessian_values"].copy()
    diag = cp[:-1]  # the diagonal entry leads each lower CSC column
    vals[diag] += 1e-3  # LM-style damping keeps the system well posed
    A = ow.csc_to_dense(n, cp, ri, vals)
    A = A + np.tril(A, -1).T
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n)
    x_true = np.linalg.solve(A, b)

    def solve_with(mixed):
        ch = capi.HipCholesky(n, cp, ri, ordering=capi.ORDERING_NESTED_DISSECTION)
        if mixed:
            ch.set_f32_updates(True)
        ch.factorize(vals)
        return ch, ch.solve(b)

    ch64, x64 = solve_with(False)
    ch32, x32 = solve_with(True)
    info = ch32.info()
    err64 = np.linalg.norm(x64 - x_true) / np.linalg.norm(x_true)
    err32 = np.linalg.norm(x32 - x_true) / np.linalg.norm(x_true)
    ev = np.linalg.eigvalsh(A)
    kappa = ev[-1] / ev[0]
    print(f"\nn={n} cond={kappa:.3g} err64={err64:.3g} err32={err32:.3g} info={info}")
    assert err64 < 1e-11
    assert err32 > 100 * err64  # the FP32 updates ran
    assert err32 < 10 * kappa * np.finfo(np.float32).eps
    # refinement against the FP64 matrix contracts the error by the factor's FP32 accuracy per
    # step (here ~1e-3) down to the FP64 factor's level
    x, errs = x32, [err32]
    for _ in range(5):
        x = x + ch32.solve(b - A @ x)
        errs.append(np.linalg.norm(x - x_true) / np.linalg.norm(x_true))
    print("refinement:", np.array2string(np.array(errs), precision=2))
    assert errs[1] < 1e-2 * errs[0]
    assert errs[-1] < max(1e-11, 10 * err64)


def deviation(run, ora):
    """Largest relative difference from the oracle's record, over new_error, new_error_linear
    and lambda, after checking the discrete trajectory (iterations, accept decisions, status;
    the best index up to ties at the converged floor)."""
    assert run["status"] == ora["status"]
    assert run["failure_reason"] == ora["failure_reason"]
    assert len(run["iterations"]) == len(ora["iterations"])
    hb, ob = run["best_index"], ora["best_index"]
    if hb != ob:
        e = [it["new_error"] for it in ora["iterations"]]
        assert e[hb] == pytest.approx(e[ob], rel=1e-12), (hb, ob, e[hb], e[ob])
    dev = 0.0
    for a, b in zip(run["iterations"], ora["iterations"]):
        assert a["iteration"] == b["iteration"]
        assert a["update_accepted"] == b["update_accepted"]
        for k in ("new_error", "new_error_linear", "current_lambda"):
            dev = max(dev, abs(a[k] - b[k]) / max(abs(b[k]), 1e-12))
    return dev


@pytest.mark.parametrize("name,gen,params", [
    ("pose_graph", lambda: capi.gen_pose_graph(2000, 150, 7), dict()),
    ("ba_circle", lambda: capi.gen_ba_circle(8, 150, 3), dict()),
    ("localization", lambda: capi.gen_localization(30, 25, 0.3, 5), dict()),
    ("pose_graph_dynamic", lambda: capi.gen_pose_graph(3000, 400, 11),
     dict(lambda_update_type=capi.LAMBDA_DYNAMIC)),
])
def test_mixed_form_follows_the_oracle(name, gen, params):
    """The mixed form keeps the oracle's discrete trajectory and is as close to its records as
    the FP64 method: within 1e-8, or within 10x the FP64 device path's own largest deviation
    where the problem is sensitive (late LM systems of the pose graphs are conditioned ~1e9,
    where a trial step's cost and the DYNAMIC lambda, a function of the gain ratio, move with
    the solve's rounding)."""
    prob = gen()
    p = capi.default_params(**params)
    opt = capi.HipOptimizer(prob, p, mixed_options())
    opt.refinement_residuals(track=1)
    hip = opt.optimize()
    f64 = capi.HipOptimizer(prob, p, capi.solver_options(
        linear_solver=capi.SOLVER_CHOLESKY, ordering=capi.ORDERING_NESTED_DISSECTION)).optimize()
    ora = ow.OracleProblem(prob).optimize(p)
    res = opt.refinement_residuals()
    stats = opt.linear_solver_stats()
    d_mixed, d_f64 = deviation(hip, ora), deviation(f64, ora)
    print(f"\n{name}: {stats}, refinement residuals max {res.max():.3g}, deviation from the "
          f"oracle: mixed {d_mixed:.3g}, FP64 path {d_f64:.3g}")
    assert d_mixed <= max(1e-8, 10 * d_f64)
    np.testing.assert_allclose(hip["values"], ora["values"], rtol=0,
                               atol=1e-7 * max(1.0, np.abs(ora["values"]).max()))
    assert len(res) == stats["solves"] == len(hip["iterations"]) - 1
    assert 1 <= stats["max_iterations"] <= 30
    assert res.max() < 1e-10


def test_pose_graph_100k_nodes_mixed():
    prob = capi.gen_pose_graph(100000, 30000, 42)
    params = capi.default_params(iterations=12)
    opt = capi.HipOptimizer(prob, params, mixed_options())
    opt.refinement_residuals(track=1)
    hip = opt.optimize()
    ora = ow.OracleProblem(prob).optimize(params, ordering=nd_permutation(prob))
    assert_same_trajectory_up_to_ties(hip, ora)
    res = opt.refinement_residuals()
    stats = opt.linear_solver_stats()
    print(f"\n100k pose graph: {stats}, refinement residuals "
          f"{np.array2string(res, precision=2)}")
    assert res.max() < 1e-10
