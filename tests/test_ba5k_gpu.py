"""The 5000-camera bundle adjustment at its full size (5000 cameras of 9 dof, 1 000 000 points,
~5M observations, Huber loss; the learned-preconditioner configuration).

* Against the oracle: tests/golden/ba5k_oracle.json holds the CPU FP64 oracle's LM run on this
  problem (reduced camera system in the nested-dissection order of the covisibility graph;
  tests/golden/make_ba5k_oracle.py). The reference LM ends FAILED with LAMBDA_OUT_OF_BOUNDS
  after 14 iterations here: most trial steps carry some points onto a camera's focal plane,
  where the Snavely projection divides by a depth near zero, and their costs reach 1e19-1e23.
  The plain FP64 device run keeps the same 15 records: the same decisions, status and best
  index; accepted costs, linear-model errors and lambda to 1e-9; rejected trial costs to 1e-4
  (a cost that divides by a near-zero depth moves by ~1e-6 relative under rounding-level
  differences of the step, and the factorization's atomic scatter orders its sums differently
  from run to run; measured 5e-8 to 4e-6).
* PCG on the reduced camera system, strict mode (the LDL^T solve whenever a solve misses its
  tolerance): block Jacobi converges on this problem's systems, and the learned correction,
  trained on systems logged from the 1000-camera problem's own solves, is measured against it
  (PCG iterations and step error against the LDL^T solve, printed); with either, the run keeps
  the oracle's decisions and final cost."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from symforce_amd import _capi as capi

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "ba5k_oracle.json"
FIXTURE_SHA256 = "b914e9de63a232c31c21688fe4f24aab9f7b6c6c5dae9f72818437f21c0b97e0"


def problem_digest(p):
    """The digest make_ba5k_oracle.py recorded for the problem it solved."""
    h = hashlib.sha256()
    for a in (p.types, p.values, p.ftypes, p.fptr, p.fargs):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def ba5k():
    return capi.gen_ba_snavely(5000, 1000000, 42)


@pytest.fixture(scope="module")
def oracle_run(ba5k):
    raw = FIXTURE.read_bytes()
    assert hashlib.sha256(raw).hexdigest() == FIXTURE_SHA256, "fixture changed"
    data = json.loads(raw)
    assert data["problem"]["sha256"] == problem_digest(ba5k), "generator output changed"
    assert data["problem"]["n_obs"] == int(ba5k.info["n_obs"])
    return data


def params():
    return capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100,
                               early_exit_min_reduction=1e-6)


def assert_follows(run, want, trial_rtol=1e-4):
    assert run["status"] == want["status"]
    assert run["failure_reason"] == want["failure_reason"]
    assert run["best_index"] == want["best_index"]
    recs = want["iterations"]
    assert len(run["iterations"]) == len(recs)
    for a, b in zip(run["iterations"], recs):
        assert a["iteration"] == b["iteration"]
        assert a["update_accepted"] == b["update_accepted"]
        assert a["current_lambda"] == pytest.approx(b["current_lambda"], rel=1e-9)
        assert a["new_error_linear"] == pytest.approx(b["new_error_linear"], rel=1e-9)
        rtol = 1e-9 if (b["update_accepted"] or b["iteration"] < 0) else trial_rtol
        assert a["new_error"] == pytest.approx(b["new_error"], rel=rtol), (a, b)


def test_ba5k_plain_follows_the_oracle_run(ba5k, oracle_run):
    opts = capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                               ordering=capi.ORDERING_NESTED_DISSECTION)
    opt = capi.HipOptimizer(ba5k, params(), opts)
    info = opt.solver_info()
    hip = opt.optimize()
    assert oracle_run["status"] == capi.STATUS_FAILED
    assert oracle_run["failure_reason"] == capi.FAILURE_LAMBDA_OUT_OF_BOUNDS
    assert_follows(hip, oracle_run)
    best = hip["iterations"][hip["best_index"]]["new_error"]
    print(f"\n5000 cameras: {info}, best cost {best!r} (oracle {oracle_run['final_cost']!r})")


@pytest.fixture(scope="module")
def trained_params():
    """The learned correction trained on systems logged from the 1000-camera problem's own
    solves (the configuration's training set)."""
    p1 = capi.gen_ba_snavely(1000, 200000, 42)
    o1 = capi.HipOptimizer(p1, capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC,
                                                   iterations=12),
                           capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                                               ordering=capi.ORDERING_NESTED_DISSECTION,
                                               use_pcg=1, pcg_max_iters=300, pcg_rel_tol=1e-6,
                                               pcg_preconditioner=capi.PRECOND_LEARNED))
    o1.log_systems(8)
    o1.optimize()
    assert o1.logged_systems() == 8
    losses = o1.train_preconditioner(300, lr=1e-3, probes=4, seed=11)
    assert losses[-50:].mean() <= losses[:50].mean()
    return o1.preconditioner_params()


@pytest.mark.parametrize("pre", [capi.PRECOND_BLOCK_JACOBI, capi.PRECOND_LEARNED])
def test_ba5k_pcg_forms(ba5k, oracle_run, trained_params, pre):
    opts = capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                               ordering=capi.ORDERING_NESTED_DISSECTION, use_pcg=2,
                               pcg_max_iters=300, pcg_rel_tol=1e-8, pcg_preconditioner=pre)
    opt = capi.HipOptimizer(ba5k, params(), opts)
    opt.preconditioner_params(trained_params)
    opt.log_systems(2)
    run = opt.optimize()
    stats = opt.linear_solver_stats()
    # strict mode: a solve that misses 1e-8 is the LDL^T solve, so the decisions are the
    # oracle's and the best cost is the oracle's final cost to 1e-6 (the time-to-cost bar)
    for a, b in zip(run["iterations"], oracle_run["iterations"]):
        assert a["update_accepted"] == b["update_accepted"]
    assert run["status"] == oracle_run["status"]
    best = run["iterations"][run["best_index"]]["new_error"]
    assert best == pytest.approx(oracle_run["final_cost"], rel=1e-6)
    report = []
    for idx in range(opt.logged_systems()):
        it_bj, err_bj = opt.eval_preconditioner(idx, 1e-6, 2000, learned=False)
        it_nn, err_nn = opt.eval_preconditioner(idx, 1e-6, 2000, learned=True)
        report.append(dict(system=idx, pcg_iters_block_jacobi=it_bj, step_err_block_jacobi=err_bj,
                           pcg_iters_learned=it_nn, step_err_learned=err_nn))
        assert 0 < it_bj < 2000 and err_bj < 1e-4  # block Jacobi converges to the LDL^T step
    name = "learned" if pre == capi.PRECOND_LEARNED else "block Jacobi"
    print(f"\n{name}: {stats}, best cost gap {(best - oracle_run['final_cost']) / best:.3g}, "
          f"systems: {report}")
