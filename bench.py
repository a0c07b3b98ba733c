"""Benchmark: Levenberg-Marquardt on the 1000-camera bundle adjustment, one MI355X.

Workload: symhip_gen_ba_snavely(1000 cameras, 200000 points, seed 42) -- Pose3 + (f, k1, k2) per
camera, ~1.0M Snavely reprojection observations with Huber loss, solved with the Schur backend
(landmark elimination + reduced camera system + supernodal LDL^T) in FP64, DYNAMIC lambda.
A "step" is one LM iteration (relinearize, damp, Schur eliminate, factor, solve, retract,
relinearize, accept/reject). Values, factor tables and all linearization storage are resident in
HBM before the timed region; the host reads back one iteration record per step.

Prints one JSON line (value = LM iterations/s of the timed FP64 run) with:
  time_to_cost  wall time until the best cost is within 1e-6 relative of the FP64 oracle's final
                cost (tests/golden/ba1k_oracle.json) for the plain and the fast forms
  roofline      the dominant kernel's achieved bandwidth (or FLOP rate) from HIP events
  cpu_baseline  the CPU FP64 oracle timed on this host for one LM iteration of the same problem
                (about 70 s on one core; the one-time symbolic analysis is included)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

from symforce_amd import _capi as capi  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "LM iters/s and time to FP64-oracle cost (rel 1e-6), 1k-cam BA; fast vs plain"
ORACLE_FIXTURE = REPO / "tests" / "golden" / "ba1k_oracle.json"
TRAFFIC_PROFILE = REPO / "profiles" / "ba1k_plain_traffic.txt"


def measured_traffic(kernel_prefix):
    """HBM bytes per launch of a kernel from the committed PMC summary (rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counters on a 512 MiB copy
    kernel); None when the summary is absent."""
    if not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith(kernel_prefix):
            return float(line.split()[-1]) * 1e6
    return None


def lm_params(iterations, early_exit=1e-6):
    return capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=iterations,
                               early_exit_min_reduction=early_exit)


def run_gpu(args):
    t0 = time.time()
    prob = capi.gen_ba_snavely(1000, 200000, 42)
    t_gen = time.time() - t0
    opts = capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                               ordering=capi.ORDERING_NESTED_DISSECTION)
    t0 = time.time()
    opt = capi.HipOptimizer(prob, lm_params(args.warmup + args.steps, early_exit=0.0), opts)
    t_setup = time.time() - t0
    info = opt.solver_info()
    # warmup (same start point, same trajectory prefix); --warmup 0 skips it
    if args.warmup > 0:
        opt.optimize_resident(num_iterations=args.warmup)
    # timed: exactly K LM iterations, early exit disabled
    capi.hip_lib()  # ensure loaded
    t0 = time.perf_counter()
    res = opt.optimize_resident(num_iterations=args.steps)
    t_run = time.perf_counter() - t0
    n_iter = len(res["iterations"]) - 1
    ms_per_step = 1e3 * t_run / n_iter
    # phase timing run (HIP events on the optimizer's stream), same K
    opt.enable_timing(True)
    opt.optimize_resident(num_iterations=args.steps)
    phases = opt.phase_times()
    opt.enable_timing(False)
    # time to the oracle's cost: plain FP64 to convergence
    target = None
    if ORACLE_FIXTURE.exists():
        target = json.loads(ORACLE_FIXTURE.read_text())["final_cost"]
    plain = fast = fast_alt = None
    if not args.no_converge:
        plain = time_to_cost(prob, opts, target)
        if not args.no_fast:
            fast = time_to_cost(prob, fast_options(args), target)
            alt = "f64" if args.fast_precision == "f32" else "f32"
            fast_alt = dict(precision=alt,
                            **time_to_cost(prob, fast_options(args, precision=alt), target))
    # the fast form's iteration rate and phases over the same K steps
    fast_ms, fast_phases, fast_lin = None, {}, None
    if not args.no_fast:
        fopt = capi.HipOptimizer(prob, lm_params(args.warmup + args.steps, early_exit=0.0),
                                 fast_options(args, switch=False))
        if args.warmup > 0:
            fopt.optimize_resident(num_iterations=args.warmup)
        t0 = time.perf_counter()
        fres = fopt.optimize_resident(num_iterations=args.steps)
        fast_ms = 1e3 * (time.perf_counter() - t0) / (len(fres["iterations"]) - 1)
        fopt.enable_timing(True)
        fopt.optimize_resident(num_iterations=args.steps)
        fast_phases = fopt.phase_times()
        fast_lin = fopt.linear_solver_stats()
    return dict(prob=prob, opt=opt, info=info, t_gen=t_gen, t_setup=t_setup,
                ms_per_step=ms_per_step, n_iter=n_iter, phases=phases, res=res, plain=plain,
                fast=fast, fast_alt=fast_alt, target=target, fast_ms=fast_ms,
                fast_phases=fast_phases,
                fast_lin=fast_lin)


def fast_options(args, switch=True, precision=None):
    """The fast form: block-Jacobi PCG on the reduced camera system (truncated at the tolerance or
    the iteration cap; with --pcg-forcing above 0 the tolerance tightens with the last accepted
    step's reduction), in the FP32 precision form by default (W, Z and S stored in FP32, FP64
    sums, cost and decisions), switching to the exact supernodal LDL^T (and FP64 storage) once an
    accepted step's relative reduction falls below --pcg-exact-below. The defaults (fixed
    tolerance 3e-2, 20 iterations, exact below 1e-4) come from a sweep on the 1000-camera
    problem: adaptive forcing and later switches let LM's early exit stop 1.3-1.5e-6 above the
    oracle's cost, this setting ends within 2e-7 of it in both storage precisions."""
    prec = args.fast_precision if precision is None else precision
    return capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                               ordering=capi.ORDERING_NESTED_DISSECTION,
                               precision=capi.PRECISION_F32_JAC if prec == "f32" else
                               capi.PRECISION_F64,
                               use_pcg=1, pcg_max_iters=args.pcg_max_iters,
                               pcg_rel_tol=args.pcg_tol,
                               pcg_exact_below_reduction=args.pcg_exact_below if switch else 0.0,
                               pcg_forcing=args.pcg_forcing)


def time_to_cost(prob, opts, target):
    """Runs one optimizer form from the initial Values to convergence and reports the wall time
    at which its best cost first came within 1e-6 relative of the oracle's final cost."""
    conv = capi.HipOptimizer(prob, lm_params(100), opts)
    t0 = time.perf_counter()
    cres = conv.optimize_resident()
    t_conv = time.perf_counter() - t0
    times = conv.iteration_times()
    lin = conv.linear_solver_stats()
    best = np.minimum.accumulate([it["new_error"] for it in cres["iterations"]])
    final_cost = float(best[-1])
    ttc = None
    if target is not None:
        hit = np.nonzero(best <= target * (1 + 1e-6))[0]
        ttc = float(times[hit[0]] / 1e3) if hit.size else None
    return dict(time_to_cost_s=ttc, final_cost=final_cost,
                iterations=len(cres["iterations"]) - 1, converge_s=t_conv,
                status=cres["status"], linear_solver=lin,
                rel_gap=None if target is None else (final_cost - target) / target)


def speedup(plain, fast):
    if not plain or not fast or not plain["time_to_cost_s"] or not fast["time_to_cost_s"]:
        return None
    return plain["time_to_cost_s"] / fast["time_to_cost_s"]


def roofline(r):
    """The largest single kernel launch of an LM iteration is the reduced-camera-system assembly
    (one launch; its phase timer brackets exactly that kernel with HIP events on the optimizer's
    stream). It reduces every off-diagonal S block over its observation pairs on the FP64
    matrix cores (K = 3 per pair, packed densely: 4 pairs per 3 MFMAs): the useful work is
    81 entries x 3 FMAs x 2 = 486 flops per pair, and the unique HBM bytes are the pair list, Z
    once and the S blocks out (the diagonal blocks' own-observation sums are reduced per camera
    in the reduced-rhs kernel). The supernodal LDL^T phase (one diagonal+panel launch and one
    update launch per level) is reported beside it as a phase: flops from the symbolic analysis
    over the phase time."""
    info = r["info"]
    ph = r["phases"]
    ms = ph["rcs_assemble"][0] / max(1, ph["rcs_assemble"][1])
    n_obs = int(r["prob"].info["n_obs"])
    flops = info["pairs"] * 486
    bytes_ = info["pairs"] * 8 + n_obs * 27 * 8 + info["blocks"] * 81 * 8
    achieved = flops / (ms * 1e-3) / 1e12
    fac_ms = ph["factor"][0] / max(1, ph["factor"][1])
    fac_tf = info["flops"] / (fac_ms * 1e-3) / 1e12
    return dict(kernel="rcs_assemble_kernel<9, double, double>", bound="mfma", achieved=achieved,
                peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=achieved / FP64_PEAK_TFLOPS,
                traffic=measured_traffic("void rcs_assemble_kernel<9, double, double>"), ms_per_launch=ms,
                algorithmic_flops=flops,
                algorithmic_bytes=bytes_,
                hbm_gbs=bytes_ / (ms * 1e-3) / 1e9,
                ldlt_factor_phase={"ms": fac_ms, "flops": info["flops"], "tflops": fac_tf,
                                   "frac_fp64": fac_tf / FP64_PEAK_TFLOPS,
                                   "levels": info["levels"]})


def batch_leg(n_problems=4096, cpu_sample=64):
    """Batched form: n independent robot-3D-localization problems (5-50 poses, 5-40 landmarks,
    30 % visible) in one persistent launch, FP64 and FP32-Jacobian forms; the CPU oracle solves
    a sample of the same problems one after another on one core."""
    probs = capi.gen_localization_batch(n_problems, seed=42)
    params = capi.default_params(iterations=50)
    out = {"workload": f"robot3d_localization_x{n_problems}", "n_problems": n_problems}
    for name, prec in (("f64", capi.PRECISION_F64), ("f32_jac", capi.PRECISION_F32_JAC)):
        b = capi.HipBatch(probs, params, precision=prec)
        b.solve(records=False)  # warm-up
        runs = [b.solve(records=False) for _ in range(3)]
        r = min(runs, key=lambda x: x["total_ms"])
        ok = sum(p["status"] == capi.STATUS_SUCCESS for p in r["problems"])
        out[name] = {"problems_per_s": n_problems / (r["total_ms"] * 1e-3),
                     "launch_ms": r["total_ms"],
                     "p50_solve_ms": float(np.percentile(r["solve_ms"], 50)),
                     "p99_solve_ms": float(np.percentile(r["solve_ms"], 99)),
                     "converged": ok}
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_wrap as ow
    t0 = time.perf_counter()
    for prob in probs[:cpu_sample]:
        ow.OracleProblem(prob).optimize(params)
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"problems_per_s": cpu_sample / dt, "cores": 1, "kind": "port",
                           "sample": f"first {cpu_sample} problems, FP64 oracle LM, sequential"}
    return out


def cpu_baseline(prob, iters=1):
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_wrap as ow

    t0 = time.time()
    op = ow.OracleProblem(prob)
    t_create = time.time() - t0
    secs, _ = ow.time_iterations(op, lm_params(iters, early_exit=0.0), iters, solver=ow.SCHUR,
                                 c_dim=int(prob.info["c_dim"]))
    per_iter = float(np.mean(secs))
    return dict(value=1.0 / per_iter, unit="LM iters/s", cores=1, kind="port",
                sample=f"CPU FP64 oracle (restated reference LM + SparseSchurSolver), same 1k-cam "
                       f"problem, {iters} LM iteration(s) after {t_create:.1f} s setup: "
                       f"{per_iter:.2f} s/iter (incl. the one-time symbolic analysis)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched small-problem leg")
    ap.add_argument("--no-fast", action="store_true", help="skip the fast-form legs")
    ap.add_argument("--fast-precision", choices=("f32", "f64"), default="f32",
                    help="storage precision of the fast form's W, Z and S")
    ap.add_argument("--pcg-tol", type=float, default=3e-2)
    ap.add_argument("--pcg-max-iters", type=int, default=20)
    ap.add_argument("--pcg-forcing", type=float, default=0.0,
                    help="inexact-Newton forcing of the PCG tolerance (0: fixed tolerance)")
    ap.add_argument("--pcg-exact-below", type=float, default=1e-4,
                    help="switch the fast form to the exact LDL^T below this LM reduction")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the run-to-convergence (time-to-cost) leg, e.g. under a profiler")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the optimizer path runs on a single GPU "
              "(landmark sharding over several GPUs is designed but not built). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if capi.hip_lib().symhip_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    r = run_gpu(args)
    roof = roofline(r)
    out = {
        "metric": METRIC,
        "value": 1e3 / r["ms_per_step"],
        "unit": "LM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, BA along a random-walk trajectory)",
        "config": {"workload": "ba_snavely_1000cams_200kpts_1Mobs_huber_schur",
                   "n_obs": int(r["prob"].info["n_obs"]), "solver": "schur+supernodal_ldlt", "ordering": "nested_dissection",
                   "lambda": "dynamic", "form": "plain_fp64"},
        "time_to_cost": {"target_cost": r["target"], "plain": r["plain"], "fast": r["fast"],
                         "fast_other_precision": r["fast_alt"],
                         "fast_form": {"solver": "schur+block_jacobi_pcg",
                                       "dtype": ("f32 W/Z/S storage, f64 sums/cost"
                                                 if args.fast_precision == "f32" else "f64"),
                                       "pcg_rel_tol": args.pcg_tol,
                                       "pcg_max_iters": args.pcg_max_iters,
                                       "pcg_forcing": args.pcg_forcing,
                                       "exact_ldlt_below_reduction": args.pcg_exact_below},
                         "speedup": speedup(r["plain"], r["fast"])},
        "setup_s": r["t_setup"],
        "phases_ms_per_iter": {k: v[0] / max(1, v[1]) for k, v in r["phases"].items() if v[1]},
        "fast_ms_per_step": r["fast_ms"],
        "fast_phases_ms_per_iter": {k: v[0] / max(1, v[1]) for k, v in r["fast_phases"].items()
                                    if v[1]} or None,
        "fast_linear_solver": r["fast_lin"],
        "roofline": roof,
    }
    if not args.no_batch:
        out["batch"] = batch_leg()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(r["prob"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
