"""Runs the CPU FP64 oracle on the 5000-camera Snavely bundle adjustment
(symhip_gen_ba_snavely(5000, 1000000, seed 42): ~5M observations, Huber loss, Schur solver,
DYNAMIC lambda, the reduced camera system eliminated in the nested-dissection order of the
camera covisibility graph) and records every iteration's statistics plus a checksum of the
input problem. The full-size GPU test of this configuration and the benchmark read the fixture.
Takes ~25 minutes on one core (the run ends after 14 iterations: the reference LM's lambda
leaves its bounds).  Run from the repo root:
    python tests/golden/make_ba5k_oracle.py
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "tests"))

from symforce_amd import _capi as capi  # noqa: E402
import oracle_wrap as ow  # noqa: E402

OUT = Path(__file__).resolve().parent / "ba5k_oracle.json"


def problem_digest(p):
    h = hashlib.sha256()
    for a in (p.types, p.values, p.ftypes, p.fptr, p.fargs):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def params():
    return capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100,
                               early_exit_min_reduction=1e-6)


def main(n_cams=5000, n_points=1000000, seed=42):
    t0 = time.time()
    p = capi.gen_ba_snavely(n_cams, n_points, seed)
    order = ow.schur_elimination_order(p)
    op = ow.OracleProblem(p)
    t1 = time.time()
    r = op.optimize(params(), solver=ow.SCHUR, c_dim=int(p.info["c_dim"]), ordering=order)
    t2 = time.time()
    best = r["iterations"][r["best_index"]]["new_error"]
    data = dict(
        problem=dict(generator="symhip_gen_ba_snavely", n_cams=n_cams, n_points=n_points,
                     seed=seed, n_obs=int(p.info["n_obs"]), sha256=problem_digest(p)),
        params=dict(lambda_update_type="DYNAMIC", iterations=100, early_exit_min_reduction=1e-6),
        solver="schur", ordering="nested dissection of the camera covisibility graph",
        status=r["status"], failure_reason=r["failure_reason"], best_index=r["best_index"],
        final_cost=best, iterations=r["iterations"], setup_seconds=t1 - t0,
        optimize_seconds=t2 - t1,
        seconds_per_iteration=(t2 - t1) / max(1, len(r["iterations"]) - 1))
    OUT.write_text(json.dumps(data, indent=1))
    print("wrote", OUT, "final cost", best, "iterations", len(r["iterations"]), "time", t2 - t1)


if __name__ == "__main__":
    main()
