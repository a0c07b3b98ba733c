"""Runs the CPU FP64 oracle to convergence on the 1000-camera Snavely bundle adjustment
(symhip_gen_ba_snavely(1000, 200000, seed 42), Huber loss, Schur solver, DYNAMIC lambda as the
reference BAL example uses) and records every iteration's statistics plus a checksum of the input
problem. The benchmark's time-to-cost target and the full-size GPU trajectory test read the
resulting fixture. Takes ~10 minutes on one core.  Run from the repo root:
    python tests/golden/make_ba1k_oracle.py
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

OUT = Path(__file__).resolve().parent / "ba1k_oracle.json"


def problem_digest(p):
    h = hashlib.sha256()
    for a in (p.types, p.values, p.ftypes, p.fptr, p.fargs):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def params():
    return capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=100)


def main(n_cams=1000, n_points=200000, seed=42):
    t0 = time.time()
    p = capi.gen_ba_snavely(n_cams, n_points, seed)
    op = ow.OracleProblem(p)
    t1 = time.time()
    r = op.optimize(params(), solver=ow.SCHUR, c_dim=int(p.info["c_dim"]))
    t2 = time.time()
    best = r["iterations"][r["best_index"]]["new_error"]
    data = dict(
        problem=dict(generator="symhip_gen_ba_snavely", n_cams=n_cams, n_points=n_points,
                     seed=seed, n_obs=int(p.info["n_obs"]), sha256=problem_digest(p)),
        params=dict(lambda_update_type="DYNAMIC", iterations=100, early_exit_min_reduction=1e-6),
        solver="schur", status=r["status"], best_index=r["best_index"], final_cost=best,
        iterations=r["iterations"], setup_seconds=t1 - t0, optimize_seconds=t2 - t1,
        seconds_per_iteration=(t2 - t1) / max(1, len(r["iterations"]) - 1))
    OUT.write_text(json.dumps(data, indent=1))
    print("wrote", OUT, "final cost", best, "iterations", len(r["iterations"]), "time", t2 - t1)


if __name__ == "__main__":
    main()
