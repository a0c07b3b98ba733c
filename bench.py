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
  roofline      the dominant work, the supernodal LDL^T factorization phase (per level a
                diagonal/panel kernel and an update kernel; the phase is the largest share of an
                iteration): FP64 flops from the symbolic analysis over the phase's HIP-event
                time; the reduced-camera-system assembly kernel is reported beside it
  cpu_baseline  the CPU FP64 oracle timed on this host's cores: 3 steady LM iterations of the
                same problem after one untimed iteration (setup and the one-time symbolic
                analysis excluded), the reduced camera system ordered by nested dissection as the
                reference orders it with METIS; also the -ffast-math build of the same sources
  batch         4096 robot-3D-localization problems in one persistent launch (FP64 and FP32
                Jacobian forms; launch autotuner; the kernel's FP64 issue rate from the committed
                PMC summary), against the CPU oracle on one core and on all cores
  pose_graph    the 100 000-node pose graph (Cholesky path): LM iterations/s, phases, the
                mixed-precision factorization form, run to convergence, the factorization
                phase's roofline, and the CPU oracle on the same problem and elimination order
  ba5k          the 5000-camera bundle adjustment (~5M observations): LM iterations/s and phases,
                the run against the CPU oracle's recorded run (tests/golden/ba5k_oracle.json),
                and strict PCG with block Jacobi and with the learned correction trained on the
                1000-camera problem's systems (iterations, fallbacks, final-cost gap)
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


def traffic_table():
    """Per-kernel rows of the committed PMC summary (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
    passes, each scaled by the same counters on a 512 MiB copy kernel): {kernel: (dispatches,
    HBM MB per launch)}; empty when the summary is absent."""
    rows = {}
    if not TRAFFIC_PROFILE.exists():
        return rows
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 5 and not line.startswith(("#", "kernel", "calibration")):
            try:
                rows[" ".join(parts[:-4])] = (int(parts[-4]), float(parts[-1]))
            except ValueError:
                continue
    return rows


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
            fast = time_to_cost(prob, fast_options(args), target, early_exit=1e-7)
            alt = "f64" if args.fast_precision == "f32" else "f32"
            fast_alt = dict(precision=alt,
                            **time_to_cost(prob, fast_options(args, precision=alt), target,
                                           early_exit=1e-7))
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


def time_to_cost(prob, opts, target, early_exit=1e-6):
    """Runs one optimizer form from the initial Values to convergence and reports the wall time
    at which its best cost first came within 1e-6 relative of the oracle's final cost. The plain
    form stops as the oracle's run did (relative reduction below 1e-6); the fast forms run to
    1e-7: their exact tail converges linearly here (Gauss-Newton on a problem with nonzero
    residuals, ~0.65 of the gap per step), so an exit at 1e-6 can stop them anywhere up to ~2e-6
    above the minimum, on either side of the 1e-6 target from run to run. The time reported is
    when the target is first reached, not when the run stops."""
    conv = capi.HipOptimizer(prob, lm_params(100, early_exit=early_exit), opts)
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
    """The dominant work of an LM iteration is the supernodal LDL^T factorization phase: per
    level of the supernodal elimination tree one diagonal kernel (diag_trsm_kernel: each
    supernode's diagonal block and its panel rows eliminated in one right-looking sweep, the
    forward solve fused in) and one update kernel (update_kernel<32>: the trailing products
    into ancestors on the FP64 matrix cores). Its algorithmic work is the symbolic analysis'
    flop count (per supernode cols^3/3 + below x cols^2 + below^2 x cols), and the phase timer
    brackets the phase with HIP events on the optimizer's stream, so achieved = flops / phase
    time against the FP64 matrix peak. traffic: the HBM bytes of the phase's kernels per LM
    iteration from the committed PMC summary.

    The reduced-camera-system assembly (one launch per iteration) is reported beside it: 486
    flops per off-diagonal observation pair on the FP64 matrix cores; unique bytes are the pair
    list, Z once and the S blocks out."""
    info = r["info"]
    ph = r["phases"]
    table = traffic_table()
    iters_in_profile = table.get("void rcs_assemble_kernel<9, double, double>", (0, 0))[0]
    fac_ms = ph["factor"][0] / max(1, ph["factor"][1])
    fac_tf = info["flops"] / (fac_ms * 1e-3) / 1e12
    fac_traffic = None
    if iters_in_profile:
        fac_traffic = sum(d * mb for k, (d, mb) in table.items()
                          if k.startswith(("void diag_trsm_kernel", "void update_kernel"))
                          ) * 1e6 / iters_in_profile
    ms = ph["rcs_assemble"][0] / max(1, ph["rcs_assemble"][1])
    n_obs = int(r["prob"].info["n_obs"])
    rcs_flops = info["pairs"] * 486
    rcs_bytes = info["pairs"] * 8 + n_obs * 27 * 8 + info["blocks"] * 81 * 8
    rcs_tf = rcs_flops / (ms * 1e-3) / 1e12
    rcs_traffic = table.get("void rcs_assemble_kernel<9, double, double>", (0, None))[1]
    rcs_traffic = None if rcs_traffic is None else rcs_traffic * 1e6
    return dict(
        kernel=(f"supernodal LDL^T factorization phase: diag_trsm_kernel + update_kernel<32>"
                f" per level, {info['levels']} levels"),
        bound="mfma", achieved=fac_tf, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
        frac=fac_tf / FP64_PEAK_TFLOPS, traffic=fac_traffic, ms_per_launch=fac_ms,
        algorithmic_flops=info["flops"],
        flops_split={"diagonal_and_panels": info["diag_panel_flops"],
                     "trailing_updates": info["update_flops"]},
        levels=info["levels"], supernodes=info["supernodes"],
        rcs_assemble={"kernel": "rcs_assemble_kernel<9, double, double>", "bound": "mfma",
                      "achieved": rcs_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": rcs_tf / FP64_PEAK_TFLOPS, "ms_per_launch": ms,
                      "algorithmic_flops": rcs_flops, "algorithmic_bytes": rcs_bytes,
                      "traffic": rcs_traffic,
                      "traffic_over_algorithmic": (None if rcs_traffic is None
                                                   else rcs_traffic / rcs_bytes),
                      "hbm_gbs": rcs_bytes / (ms * 1e-3) / 1e9})


def host_cpu():
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def batch_flops_table():
    """Issued FP64 lane-flops per launch of each batched-kernel variant, from the committed PMC
    summary (profiles/batch_localization_f64_ops.txt)."""
    path = REPO / "profiles" / "batch_localization_f64_ops.txt"
    rows = {}
    if path.exists():
        for line in path.read_text().splitlines():
            if line.startswith("batch_lm_kernel"):
                name, rest = line[:30].strip(), line[30:].split()
                rows[name] = float(rest[-1])
    return rows


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
        chosen, config_ms = b.tune(reps=2)
        runs = [b.solve(records=False) for _ in range(3)]
        r = min(runs, key=lambda x: x["total_ms"])
        ok = sum(p["status"] == capi.STATUS_SUCCESS for p in r["problems"])
        kernel = (f"batch_lm_kernel<{'double' if prec == capi.PRECISION_F64 else 'float'}, "
                  f"{2 if chosen & 1 else 1}>")
        flops = batch_flops_table().get(kernel)
        out[name] = {"problems_per_s": n_problems / (r["total_ms"] * 1e-3),
                     "launch_ms": r["total_ms"],
                     "p50_solve_ms": float(np.percentile(r["solve_ms"], 50)),
                     "p99_solve_ms": float(np.percentile(r["solve_ms"], 99)),
                     "converged": ok,
                     "autotune": {"chosen": capi.HipBatch.BATCH_CONFIGS[chosen],
                                  "config_ms": {capi.HipBatch.BATCH_CONFIGS[c]: float(t)
                                                for c, t in enumerate(config_ms)}},
                     # FP64 VALU issue (masked lanes included) against the FP64 vector peak:
                     # the kernel is latency-bound (per-problem dependent chains)
                     "roofline": None if flops is None else {
                         "kernel": kernel, "bound": "latency",
                         "achieved": flops / (r["total_ms"] * 1e-3) / 1e12,
                         "peak": 78.6, "unit": "TFLOP/s (FP64 vector)",
                         "frac": flops / (r["total_ms"] * 1e-3) / 1e12 / 78.6,
                         "issued_fp64_lane_flops_per_launch": flops}}
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_wrap as ow
    from concurrent.futures import ThreadPoolExecutor

    def solve_cpu(prob):
        ow.OracleProblem(prob).optimize(params)

    model, cores = host_cpu()
    t0 = time.perf_counter()
    for prob in probs[:cpu_sample]:
        solve_cpu(prob)
    dt1 = time.perf_counter() - t0
    # one optimizer per thread, as the reference allows (opt/optimizer.h:23); the oracle's calls
    # release the interpreter lock, so the threads run on all cores
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(solve_cpu, probs))
    dtn = time.perf_counter() - t0
    out["cpu_baseline"] = {"problems_per_s": cpu_sample / dt1, "cores": 1, "kind": "port",
                           "sample": f"first {cpu_sample} problems, FP64 oracle LM, sequential",
                           "all_cores": {"problems_per_s": n_problems / dtn, "cores": cores,
                                         "sample": f"all {n_problems} problems, one oracle "
                                                   f"optimizer per thread"},
                           "cpu": model}
    return out


def mixed_precision_pose_graph(prob, steps):
    """The mixed-precision form of the pose graph's factorization (FP32 matrix-core trailing
    updates, FP64 diagonal blocks and solves, CG refinement against the FP64 Hessian): LM
    iterations timed and phase-timed like the FP64 form."""
    mopt = capi.HipOptimizer(prob, lm_params(steps + 2, early_exit=0.0), capi.solver_options(
        linear_solver=capi.SOLVER_CHOLESKY, ordering=capi.ORDERING_NESTED_DISSECTION,
        precision=capi.PRECISION_F32_UPDATES))
    mopt.set_values(prob.values)
    mopt.optimize_resident(num_iterations=2)
    t0 = time.perf_counter()
    mopt.optimize_resident(num_iterations=steps)
    ms = 1e3 * (time.perf_counter() - t0) / steps
    mopt.enable_timing(True)
    mopt.optimize_resident(num_iterations=steps)
    return {"ms_per_iter": ms,
            "phases_ms_per_iter": {k: v[0] / max(1, v[1])
                                   for k, v in mopt.phase_times().items() if v[1]},
            "linear_solver": mopt.linear_solver_stats(),
            "form": "FP32 matrix-core trailing updates, FP64 diagonal blocks and solves, CG "
                    "refinement against the FP64 Hessian"}


def pose_graph_leg(steps=10, cpu_iters=5, cpu=True, mixed=True):
    """Pose-graph SLAM at 100 000 Pose3 nodes (115 037 between factors, 15 037 of them Cauchy
    loop closures): the sparse Cholesky path (supernodal LDL^T over the nested-dissection
    order) timed for `steps` LM iterations with phase timing, the factorization phase against
    the FP64 matrix peak, a run to convergence, and the CPU oracle (restated reference LM +
    SparseCholeskySolver, the same elimination order) for `cpu_iters` steady iterations."""
    prob = capi.gen_pose_graph(100000, 30000, 42)
    opts = capi.solver_options(linear_solver=capi.SOLVER_CHOLESKY,
                               ordering=capi.ORDERING_NESTED_DISSECTION)
    opt = capi.HipOptimizer(prob, lm_params(steps + 2, early_exit=0.0), opts)
    info = opt.solver_info()
    opt.set_values(prob.values)
    opt.optimize_resident(num_iterations=2)
    t0 = time.perf_counter()
    opt.optimize_resident(num_iterations=steps)
    ms = 1e3 * (time.perf_counter() - t0) / steps
    opt.enable_timing(True)
    opt.optimize_resident(num_iterations=steps)
    phases = {k: v[0] / max(1, v[1]) for k, v in opt.phase_times().items() if v[1]}
    opt.enable_timing(False)
    # the mixed-precision form of the factorization (FP32 matrix-core updates, CG refinement)
    mixed_out = mixed and mixed_precision_pose_graph(prob, steps)
    conv = capi.HipOptimizer(prob, capi.default_params(iterations=100), opts)
    conv.set_values(prob.values)
    conv.optimize_resident()  # warm
    t0 = time.perf_counter()
    res = conv.optimize_resident()
    t_conv = time.perf_counter() - t0
    flops = float(info["flops"])
    out = {"workload": "pose_graph_100k_nodes_115k_edges_cauchy_loops",
           "n_loops": int(prob.info["n_loops"]), "unknowns": int(info["n"]),
           "lm_iters_per_s": 1e3 / ms, "ms_per_iter": ms, "phases_ms_per_iter": phases,
           "mixed_precision": mixed_out or None,
           "converge": {"s": t_conv, "iterations": len(res["iterations"]) - 1,
                        "status": res["status"],
                        "final_cost": res["iterations"][res["best_index"]]["new_error"]},
           "factor_roofline": {"kernel": "supernodal LDL^T factorization phase",
                               "bound": "mfma", "levels": int(info["levels"]),
                               "supernodes": int(info["supernodes"]),
                               "algorithmic_flops": flops,
                               "achieved": flops / (phases["factor"] * 1e-3) / 1e12,
                               "peak": 78.6, "unit": "TFLOP/s",
                               "frac": flops / (phases["factor"] * 1e-3) / 1e12 / 78.6}}
    if not cpu:
        return out
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_wrap as ow
    hp = capi.HipProblem(prob)
    cp, ri = hp.hessian_structure()
    _, offsets = hp.state_keys()
    node_ptr = np.append(offsets, len(cp) - 1).astype(np.int32)
    order = capi.HipCholesky(len(cp) - 1, cp, ri, node_ptr=node_ptr,
                             ordering=capi.ORDERING_NESTED_DISSECTION).inv_permutation()
    op = ow.OracleProblem(prob)
    secs, _ = ow.time_iterations(op, lm_params(cpu_iters + 1, early_exit=0.0), cpu_iters + 1,
                                 solver=ow.CHOLESKY, ordering=order)
    model, _ = host_cpu()
    out["cpu_baseline"] = {"value": 1.0 / float(np.mean(secs[1:])), "unit": "LM iters/s",
                           "cores": 1, "kind": "port", "cpu": model,
                           "sample": (f"CPU FP64 oracle (restated reference LM + "
                                      f"SparseCholeskySolver, the same nested-dissection order), "
                                      f"same problem: {cpu_iters} steady LM iterations after one "
                                      f"untimed one ({secs[0]:.1f} s with the symbolic "
                                      f"analysis): {np.mean(secs[1:]):.2f} s/iter")}
    return out


def ba5k_leg(train_prob, steps=10):
    """Bundle adjustment at 5000 cameras (1 000 000 points, ~5M observations, Huber): the plain
    FP64 Schur path timed for `steps` LM iterations with phase timing and run to its end against
    the CPU oracle's recorded run (tests/golden/ba5k_oracle.json: the reference LM ends FAILED,
    lambda out of bounds, after 14 iterations), then PCG on the reduced camera system with block
    Jacobi and with the learned correction trained on systems logged from `train_prob` (the
    1000-camera problem), strict (the LDL^T solve when a solve misses 1e-8): wall time, PCG
    iterations, fallbacks and the final-cost gap to the oracle."""
    prob = capi.gen_ba_snavely(5000, 1000000, 42)
    exact = capi.solver_options(linear_solver=capi.SOLVER_SCHUR,
                                ordering=capi.ORDERING_NESTED_DISSECTION)
    fixture = REPO / "tests" / "golden" / "ba5k_oracle.json"
    target = json.loads(fixture.read_text())["final_cost"] if fixture.exists() else None
    t0 = time.time()
    opt = capi.HipOptimizer(prob, lm_params(steps + 2, early_exit=0.0), exact)
    t_setup = time.time() - t0
    info = opt.solver_info()
    opt.set_values(prob.values)
    opt.optimize_resident(num_iterations=2)
    t0 = time.perf_counter()
    opt.optimize_resident(num_iterations=steps)
    ms = 1e3 * (time.perf_counter() - t0) / steps
    opt.enable_timing(True)
    opt.optimize_resident(num_iterations=steps)
    phases = {k: v[0] / max(1, v[1]) for k, v in opt.phase_times().items() if v[1]}
    del opt

    def run(opts, params=None):
        o = capi.HipOptimizer(prob, lm_params(100), opts)
        if params is not None:
            o.preconditioner_params(params)
        t0 = time.perf_counter()
        res = o.optimize_resident()
        dt = time.perf_counter() - t0
        best = res["iterations"][res["best_index"]]["new_error"]
        return {"s": dt, "iterations": len(res["iterations"]) - 1, "status": res["status"],
                "final_cost": best, "linear_solver": o.linear_solver_stats(),
                "rel_gap": None if target is None else (best - target) / target}

    out = {"workload": "ba_snavely_5000cams_1Mpts_5Mobs_huber_schur",
           "n_obs": int(prob.info["n_obs"]), "unknowns_rcs": int(info["n"]),
           "setup_s": t_setup, "ms_per_iter": ms, "lm_iters_per_s": 1e3 / ms,
           "phases_ms_per_iter": phases, "oracle_final_cost": target,
           "plain": run(exact)}
    trainer = capi.HipOptimizer(train_prob, capi.default_params(
        lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=12), capi.solver_options(
        linear_solver=capi.SOLVER_SCHUR, ordering=capi.ORDERING_NESTED_DISSECTION, use_pcg=1,
        pcg_max_iters=300, pcg_rel_tol=1e-6, pcg_preconditioner=capi.PRECOND_LEARNED))
    trainer.log_systems(8)
    trainer.optimize()
    t0 = time.perf_counter()
    losses = trainer.train_preconditioner(300, lr=1e-3, probes=4, seed=11)
    t_train = time.perf_counter() - t0
    params = trainer.preconditioner_params()
    del trainer
    for name, pre in (("pcg_block_jacobi", capi.PRECOND_BLOCK_JACOBI),
                      ("pcg_learned", capi.PRECOND_LEARNED)):
        out[name] = run(capi.solver_options(
            linear_solver=capi.SOLVER_SCHUR, ordering=capi.ORDERING_NESTED_DISSECTION, use_pcg=2,
            pcg_max_iters=300, pcg_rel_tol=1e-8, pcg_preconditioner=pre), params)
    out["pcg_learned"]["training"] = {"systems": 8, "steps": 300, "s": t_train,
                                      "loss_first": float(losses[0]),
                                      "loss_last": float(losses[-1]),
                                      "source": "1000-camera problem's own solves"}
    return out


def cpu_baseline(prob, iters=3):
    """The CPU FP64 oracle (the restated reference LM + SparseSchurSolver) on this host: one
    untimed LM iteration (it includes the one-time symbolic analysis), then `iters` timed ones,
    single-threaded as the reference path is; then the same with the -ffast-math build."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_wrap as ow

    order = ow.schur_elimination_order(prob)
    model, _ = host_cpu()
    res = {}
    for name, fast in (("plain", False), ("fast_math", True)):
        t0 = time.time()
        op = ow.OracleProblem(prob, fast_math=fast)
        t_create = time.time() - t0
        secs, _ = ow.time_iterations(op, lm_params(iters + 1, early_exit=0.0), iters + 1,
                                     solver=ow.SCHUR, c_dim=int(prob.info["c_dim"]),
                                     ordering=order)
        del op
        res[name] = dict(setup_s=t_create, first_iteration_s=float(secs[0]),
                         s_per_iteration=float(np.mean(secs[1:])))
    per_iter = res["plain"]["s_per_iteration"]
    return dict(value=1.0 / per_iter, unit="LM iters/s", cores=1, kind="port",
                sample=(f"CPU FP64 oracle (restated reference LM + SparseSchurSolver, S ordered "
                        f"by nested dissection), same 1k-cam problem: {iters} steady LM "
                        f"iterations after one untimed one (setup "
                        f"{res['plain']['setup_s']:.1f} s and first iteration "
                        f"{res['plain']['first_iteration_s']:.1f} s with the symbolic analysis, "
                        f"both excluded): {per_iter:.2f} s/iter"),
                cpu=model,
                fast_math={"value": 1.0 / res["fast_math"]["s_per_iteration"],
                           "s_per_iteration": res["fast_math"]["s_per_iteration"],
                           "build": "-O3 -march=native -ffast-math"},
                detail=res)


def calibrate_copy(mib=512, launches=3):
    """A device-to-device copy of known size under the PMC counters: bytes read = bytes written
    = mib MiB per launch. Uses torch's copy kernel (torch is here only as the copy's harness)."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(launches):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_mib": mib, "launches": launches}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched small-problem leg")
    ap.add_argument("--no-ba5k", action="store_true", help="skip the 5000-camera leg")
    ap.add_argument("--no-mixed", action="store_true",
                    help="pose-graph leg without the mixed-precision factorization form")
    ap.add_argument("--no-fast", action="store_true", help="skip the fast-form legs")
    ap.add_argument("--no-pose-graph", action="store_true",
                    help="skip the 100k-node pose-graph leg")
    ap.add_argument("--pose-graph-only", action="store_true",
                    help="only run the pose-graph leg's GPU part and print it (for profiling)")
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
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 512 MiB device to device three times (a kernel of known HBM "
                         "traffic, to scale the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.pose_graph_only:
        print(json.dumps(pose_graph_leg(cpu=False, mixed=not args.no_mixed)))
        return
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
    if not args.no_pose_graph:
        out["pose_graph"] = pose_graph_leg()
    if not args.no_ba5k:
        out["ba5k"] = ba5k_leg(r["prob"])
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(r["prob"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
