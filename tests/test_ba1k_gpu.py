"""The 1000-camera bundle adjustment at its full size (1000 cameras, 200 000 points, ~1M
observations; the benchmark's workload).

* Against the oracle: tests/golden/ba1k_oracle.json holds the CPU FP64 oracle's whole LM run on
  this problem (79 records, 45 minutes on one core; tests/golden/make_ba1k_oracle.py). The plain
  FP64 device run to convergence must follow its first iterations to 1e-8 relative in new_error
  with the same accept decisions and land on its final cost to 1e-9; the fast form (FP32
  storage, block-Jacobi PCG, exact LDL^T late) must reach it (its cost at most 1e-6 above the
  oracle's final cost; run to a 1e-7 reduction it goes below it). The oracle orders the
  reduced camera system naturally and the device by nested dissection, so the two runs differ
  only by rounding in the factorization.
* A size-independent property: the exact Schur step (reduced camera system assembled on the
  matrix cores, supernodal LDL^T of the 9000 x 9000 system with nested dissection: 38 levels,
  both update-tile sizes, the fused forward solve) is compared with an independent solver of the
  same reduced system, strict FP64 block-Jacobi PCG run to a relative residual of 1e-12 (strict
  mode: no step may fall back to the LDL^T). The two steps differ by cond(S) times the PCG
  residual, so the costs along the first LM iterations agree to 1e-6 relative; a factorization
  error at this size (a wrong tile, a missed update, a race between levels) moves a step by far
  more."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from symforce_amd import _capi as capi

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "ba1k_oracle.json"
FIXTURE_SHA256 = "9c9e0ec0a1fc340aef80759acd912655a078ba99dbbf0e2fd2b816178b64d08d"


def problem_digest(p):
    """The digest make_ba1k_oracle.py recorded for the problem it solved."""
    h = hashlib.sha256()
    for a in (p.types, p.values, p.ftypes, p.fptr, p.fargs):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def ba1k():
    return capi.gen_ba_snavely(1000, 200000, 42)


@pytest.fixture(scope="module")
def oracle_run(ba1k):
    raw = FIXTURE.read_bytes()
    assert hashlib.sha256(raw).hexdigest() == FIXTURE_SHA256, "fixture changed"
    data = json.loads(raw)
    assert data["problem"]["sha256"] == problem_digest(ba1k), "generator output changed"
    assert data["problem"]["n_obs"] == int(ba1k.info["n_obs"])
    return data


def _params():
    return capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100,
                               early_exit_min_reduction=1e-6)


def _best(r):
    return r["iterations"][r["best_index"]]["new_error"]


def test_ba1k_plain_follows_the_oracle_run(ba1k, oracle_run):
    opts = capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                               ordering=capi.ORDERING_NESTED_DISSECTION)
    hip = capi.HipOptimizer(ba1k, _params(), opts).optimize()
    want = oracle_run["iterations"]
    assert hip["iterations"][0]["new_error"] == pytest.approx(want[0]["new_error"], rel=1e-12)
    for a, b in zip(hip["iterations"][1:11], want[1:11]):
        assert a["iteration"] == b["iteration"]
        assert a["update_accepted"] == b["update_accepted"]
        assert a["new_error"] == pytest.approx(b["new_error"], rel=1e-8)
        assert a["new_error_linear"] == pytest.approx(b["new_error_linear"], rel=1e-8)
        assert a["current_lambda"] == pytest.approx(b["current_lambda"], rel=1e-8)
    assert hip["status"] == oracle_run["status"] == capi.STATUS_SUCCESS
    assert _best(hip) == pytest.approx(oracle_run["final_cost"], rel=1e-9)
    print(f"\nplain: {len(hip['iterations'])} records (oracle {len(want)}), final cost "
          f"{_best(hip)!r} vs {oracle_run['final_cost']!r}")


@pytest.mark.parametrize("precision", [capi.PRECISION_F32_JAC, capi.PRECISION_F64])
def test_ba1k_fast_form_reaches_the_oracle_cost(ba1k, oracle_run, precision):
    opts = capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                               ordering=capi.ORDERING_NESTED_DISSECTION, precision=precision,
                               use_pcg=1, pcg_max_iters=20, pcg_rel_tol=3e-2,
                               pcg_exact_below_reduction=1e-4)
    # run to a relative reduction of 1e-7: after the switch to exact steps the fast form converges
    # linearly, so the 1e-6 exit can stop it up to ~2e-6 above the oracle's cost. The bar is
    # reaching the oracle's cost (one-sided): the oracle's own run ended at its 1e-6 exit, and run
    # further the form goes below it (the valley keeps falling at ~2e-7 per iteration)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100,
                                 early_exit_min_reduction=1e-7)
    hip = capi.HipOptimizer(ba1k, params, opts).optimize()
    assert hip["status"] in (capi.STATUS_SUCCESS, capi.STATUS_HIT_ITERATION_LIMIT)
    its = hip["iterations"]
    print(f"\nfast form ({precision}): {len(its) - 1} iterations, best {_best(hip)!r}, last "
          f"relative reductions {[round(it['relative_reduction'], 9) for it in its[-6:]]}")
    assert _best(hip) <= oracle_run["final_cost"] * (1 + 1e-6)


def test_ba1k_exact_step_matches_tight_pcg(ba1k):
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=4,
                                 early_exit_min_reduction=0.0)
    base = dict(linear_solver=capi.SOLVER_SCHUR, ordering=capi.ORDERING_NESTED_DISSECTION)
    direct = capi.HipOptimizer(ba1k, params, capi.solver_options(**base)).optimize()
    opt = capi.HipOptimizer(ba1k, params, capi.solver_options(
        **base, use_pcg=2, pcg_max_iters=20000, pcg_rel_tol=1e-12))
    pcg = opt.optimize()
    st = opt.linear_solver_stats()
    assert st["solves"] == len(pcg["iterations"]) - 1 and st["fallbacks"] == 0
    assert len(direct["iterations"]) == len(pcg["iterations"]) == 5
    assert direct["iterations"][0]["new_error"] == pcg["iterations"][0]["new_error"]
    for d, p in zip(direct["iterations"][1:], pcg["iterations"][1:]):
        assert abs(d["new_error"] - p["new_error"]) <= 1e-6 * d["new_error"]
        assert d["update_accepted"] == p["update_accepted"]
