"""ctypes mirror of the C ABI declared in include/symhip_types.h, include/symhip.h and
include/symhip_problems.h.

The shared libraries are built in-tree into symforce_amd/_build/ (see symforce_amd/build.py).
Loading the HIP library fails loudly when it has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
BUILD_DIR = PKG_DIR / "_build"
LIB_PATH = BUILD_DIR / "libsymhip.so"
PROBLEMS_LIB_PATH = BUILD_DIR / "libsymhip_problems.so"

# ---- enums (include/symhip_types.h) ---------------------------------------------------------
TYPE_SCALAR, TYPE_ROT3, TYPE_POSE3 = 1, 3, 5
TYPE_VECTOR2, TYPE_VECTOR3, TYPE_VECTOR4, TYPE_VECTOR5, TYPE_VECTOR6 = 12, 13, 14, 15, 16
TYPE_MATRIX33, TYPE_MATRIX66 = 37, 64

FACTOR_PRIOR_POSE3 = 1
FACTOR_BETWEEN_POSE3 = 2
FACTOR_PRIOR_ROT3 = 3
FACTOR_BETWEEN_ROT3 = 4
FACTOR_SNAVELY = 5
FACTOR_SNAVELY_HUBER = 6
FACTOR_REPROJ_CAUCHY = 7
FACTOR_MATCHING = 8
FACTOR_ODOMETRY = 9
FACTOR_BETWEEN_POSE3_CAUCHY = 10
FACTOR_BARRON_VECTOR5 = 11

LAMBDA_STATIC, LAMBDA_DYNAMIC = 1, 2
STATUS_INVALID, STATUS_SUCCESS, STATUS_HIT_ITERATION_LIMIT, STATUS_FAILED = 0, 1, 2, 3
FAILURE_INVALID, FAILURE_LAMBDA_OUT_OF_BOUNDS, FAILURE_INITIAL_ERROR_NOT_FINITE = 0, 1, 2

STORAGE_DIM = {1: 1, 3: 4, 5: 7, 12: 2, 13: 3, 14: 4, 15: 5, 16: 6, 37: 9, 64: 36}
TANGENT_DIM = {1: 1, 3: 3, 5: 6, 12: 2, 13: 3, 14: 4, 15: 5, 16: 6, 37: 9, 64: 36}

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
I8P = C.POINTER(C.c_int8)
F64P = C.POINTER(C.c_double)
F32P = C.POINTER(C.c_float)


class ProblemDesc(C.Structure):
    _fields_ = [
        ("n_keys", C.c_int32),
        ("key_letters", I8P),
        ("key_subs", I64P),
        ("key_supers", I64P),
        ("key_types", I32P),
        ("values", F64P),
        ("n_factors", C.c_int32),
        ("factor_types", I32P),
        ("factor_arg_ptr", I32P),
        ("factor_args", I32P),
        ("n_opt_keys", C.c_int32),
        ("opt_keys", I32P),
    ]


class OptimizerParams(C.Structure):
    _fields_ = [
        ("verbose", C.c_int32),
        ("debug_stats", C.c_int32),
        ("check_derivatives", C.c_int32),
        ("include_jacobians", C.c_int32),
        ("debug_checks", C.c_int32),
        ("initial_lambda", C.c_double),
        ("lambda_lower_bound", C.c_double),
        ("lambda_upper_bound", C.c_double),
        ("lambda_update_type", C.c_int32),
        ("lambda_up_factor", C.c_double),
        ("lambda_down_factor", C.c_double),
        ("dynamic_lambda_update_beta", C.c_double),
        ("dynamic_lambda_update_gamma", C.c_double),
        ("dynamic_lambda_update_p", C.c_int32),
        ("use_diagonal_damping", C.c_int32),
        ("use_unit_damping", C.c_int32),
        ("keep_max_diagonal_damping", C.c_int32),
        ("diagonal_damping_min", C.c_double),
        ("iterations", C.c_int32),
        ("early_exit_min_reduction", C.c_double),
        ("early_exit_min_absolute_error", C.c_double),
        ("enable_bold_updates", C.c_int32),
    ]


def default_params(**overrides) -> OptimizerParams:
    """DefaultOptimizerParams() (reference symforce/opt/optimizer.cc:8-56)."""
    p = OptimizerParams(
        verbose=0, debug_stats=0, check_derivatives=0, include_jacobians=0, debug_checks=0,
        initial_lambda=1.0, lambda_lower_bound=0.0, lambda_upper_bound=1000000.0,
        lambda_update_type=LAMBDA_STATIC, lambda_up_factor=4.0, lambda_down_factor=0.25,
        dynamic_lambda_update_beta=2.0, dynamic_lambda_update_gamma=3.0,
        dynamic_lambda_update_p=3, use_diagonal_damping=0, use_unit_damping=1,
        keep_max_diagonal_damping=0, diagonal_damping_min=1e-6, iterations=50,
        early_exit_min_reduction=1e-6, early_exit_min_absolute_error=0.0,
        enable_bold_updates=0)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


class IterationStats(C.Structure):
    _fields_ = [
        ("iteration", C.c_int32),
        ("update_accepted", C.c_int32),
        ("current_lambda", C.c_double),
        ("new_error_linear", C.c_double),
        ("new_error", C.c_double),
        ("relative_reduction", C.c_double),
        ("update_angle_change", C.c_double),
    ]


class GncParams(C.Structure):
    """optimizer_gnc_params_t (lcmtypes/symforce.lcm:203-227)."""
    _fields_ = [
        ("gnc_update_min_reduction", C.c_double),
        ("mu_initial", C.c_double),
        ("mu_step", C.c_double),
        ("mu_max", C.c_double),
    ]


class OptimizationStats(C.Structure):
    _fields_ = [
        ("num_iterations", C.c_int32),
        ("best_index", C.c_int32),
        ("status", C.c_int32),
        ("failure_reason", C.c_int32),
    ]


def iteration_records(arr, n):
    out = []
    for k in range(n):
        r = arr[k]
        out.append(dict(iteration=r.iteration, update_accepted=bool(r.update_accepted),
                        current_lambda=r.current_lambda, new_error_linear=r.new_error_linear,
                        new_error=r.new_error, relative_reduction=r.relative_reduction,
                        update_angle_change=r.update_angle_change))
    return out


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ---- problem description held by Python (owns its arrays) ----------------------------------
class Problem:
    """A problem description (keys, values, factors) as numpy arrays plus the ProblemDesc view
    that is handed to the C ABI."""

    def __init__(self, letters, subs, supers, types, values, ftypes, fptr, fargs, opt_keys=None,
                 info=None):
        self.letters = np.ascontiguousarray(letters, dtype=np.int8)
        self.subs = np.ascontiguousarray(subs, dtype=np.int64)
        self.supers = np.ascontiguousarray(supers, dtype=np.int64)
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.ftypes = np.ascontiguousarray(ftypes, dtype=np.int32)
        self.fptr = np.ascontiguousarray(fptr, dtype=np.int32)
        self.fargs = np.ascontiguousarray(fargs, dtype=np.int32)
        self.opt_keys = (np.ascontiguousarray(opt_keys, dtype=np.int32)
                         if opt_keys is not None and len(opt_keys) else None)
        self.info = dict(info or {})
        sd = np.array([STORAGE_DIM[int(t)] for t in self.types], dtype=np.int64)
        self.storage_offsets = np.concatenate([[0], np.cumsum(sd)[:-1]]).astype(np.int64)
        assert self.values.size == int(sd.sum()), "values size does not match key types"

    @property
    def n_keys(self):
        return self.types.size

    @property
    def n_factors(self):
        return self.ftypes.size

    def desc(self, values: np.ndarray | None = None) -> ProblemDesc:
        vals = self.values if values is None else np.ascontiguousarray(values, dtype=np.float64)
        self._desc_values = vals  # keep alive
        d = ProblemDesc()
        d.n_keys = self.n_keys
        d.key_letters = ptr(self.letters, C.c_int8)
        d.key_subs = ptr(self.subs, C.c_int64)
        d.key_supers = ptr(self.supers, C.c_int64)
        d.key_types = ptr(self.types, C.c_int32)
        d.values = ptr(vals, C.c_double)
        d.n_factors = self.n_factors
        d.factor_types = ptr(self.ftypes, C.c_int32)
        d.factor_arg_ptr = ptr(self.fptr, C.c_int32)
        d.factor_args = ptr(self.fargs, C.c_int32)
        if self.opt_keys is not None:
            d.n_opt_keys = self.opt_keys.size
            d.opt_keys = ptr(self.opt_keys, C.c_int32)
        else:
            d.n_opt_keys = 0
            d.opt_keys = None
        return d


# ---- generators -------------------------------------------------------------------------------
_problems_lib = None


def problems_lib():
    global _problems_lib
    if _problems_lib is None:
        if not PROBLEMS_LIB_PATH.exists():
            raise RuntimeError(f"{PROBLEMS_LIB_PATH} is not built; run symforce_amd.build")
        lib = C.CDLL(str(PROBLEMS_LIB_PATH))
        vp = C.c_void_p
        for name, args in [
            ("symhip_gen_pose_smoothing", [C.c_int32, C.c_uint32]),
            ("symhip_gen_rot_smoothing", [C.c_int32, C.c_uint32]),
            ("symhip_gen_rot_frozen", [C.c_int32, C.c_uint32]),
            ("symhip_gen_ba_circle", [C.c_int32, C.c_int32, C.c_uint32]),
            ("symhip_gen_ba_snavely", [C.c_int32, C.c_int32, C.c_uint32]),
            ("symhip_load_bal", [C.c_char_p, C.c_double]),
            ("symhip_gen_pose_graph", [C.c_int32, C.c_int32, C.c_uint32]),
            ("symhip_gen_localization", [C.c_int32, C.c_int32, C.c_double, C.c_uint32]),
        ]:
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = vp
        lib.symhip_gen_free.argtypes = [vp]
        lib.symhip_gen_desc.argtypes = [vp, C.POINTER(ProblemDesc)]
        lib.symhip_gen_info.argtypes = [vp, C.c_char_p]
        lib.symhip_gen_info.restype = C.c_double
        lib.symhip_gen_storage_size.argtypes = [vp]
        lib.symhip_gen_storage_size.restype = C.c_int64
        lib.symhip_save_bal.argtypes = [C.POINTER(ProblemDesc), C.c_char_p]
        lib.symhip_save_bal.restype = C.c_int
        _problems_lib = lib
    return _problems_lib


_INFO_NAMES = ("epsilon", "c_dim", "n_cams", "n_points", "n_obs", "n_loops", "n_outliers",
               "track_alpha")


def _from_generated(handle) -> Problem:
    lib = problems_lib()
    d = ProblemDesc()
    lib.symhip_gen_desc(handle, C.byref(d))
    nk, nf = d.n_keys, d.n_factors
    ns = lib.symhip_gen_storage_size(handle)
    arr = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)  # noqa
    fptr = arr(d.factor_arg_ptr, nf + 1, np.int32)
    info = {}
    for name in _INFO_NAMES:
        v = lib.symhip_gen_info(handle, name.encode())
        if v == v:
            info[name] = v
    prob = Problem(
        arr(d.key_letters, nk, np.int8), arr(d.key_subs, nk, np.int64),
        arr(d.key_supers, nk, np.int64), arr(d.key_types, nk, np.int32),
        arr(d.values, ns, np.float64), arr(d.factor_types, nf, np.int32), fptr,
        arr(d.factor_args, int(fptr[-1]), np.int32),
        arr(d.opt_keys, d.n_opt_keys, np.int32) if d.n_opt_keys > 0 else None, info)
    lib.symhip_gen_free(handle)
    return prob


def gen_pose_smoothing(num_keys=10, seed=42):
    return _from_generated(problems_lib().symhip_gen_pose_smoothing(num_keys, seed))


def gen_rot_smoothing(num_keys=10, seed=42):
    return _from_generated(problems_lib().symhip_gen_rot_smoothing(num_keys, seed))


def gen_rot_frozen(num_keys=3, seed=42):
    return _from_generated(problems_lib().symhip_gen_rot_frozen(num_keys, seed))


def gen_ba_circle(n_cams=20, n_points=1000, seed=42):
    return _from_generated(problems_lib().symhip_gen_ba_circle(n_cams, n_points, seed))


def load_bal(path, huber_k=0.0):
    """Bundle Adjustment in the Large text file -> Problem (the reference example's
    ReadProblem, bundle_adjustment_in_the_large.cc:60-118)."""
    h = problems_lib().symhip_load_bal(str(path).encode(), float(huber_k))
    if not h:
        raise RuntimeError(f"could not read BAL problem from {path}")
    return _from_generated(h)


def save_bal(prob, path):
    """Writes the Snavely part of a problem as BAL text."""
    d = prob.desc()
    if problems_lib().symhip_save_bal(C.byref(d), str(path).encode()) != 0:
        raise RuntimeError(f"could not write BAL problem to {path}")


def gen_ba_snavely(n_cams=1000, n_points=200000, seed=42):
    return _from_generated(problems_lib().symhip_gen_ba_snavely(n_cams, n_points, seed))


def gen_pose_graph(n_nodes=100000, n_loops=30000, seed=42):
    return _from_generated(problems_lib().symhip_gen_pose_graph(n_nodes, n_loops, seed))


def gen_localization(n_poses=20, n_landmarks=20, visible_fraction=0.3, seed=42):
    return _from_generated(
        problems_lib().symhip_gen_localization(n_poses, n_landmarks, visible_fraction, seed))


# ---- the HIP library -----------------------------------------------------------------------
_hip_lib = None


def hip_lib():
    """Loads libsymhip.so; raises if it has not been built (no fallback exists)."""
    global _hip_lib
    if _hip_lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m symforce_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.symhip_last_error.restype = C.c_char_p
        L.symhip_device_count.restype = C.c_int
        L.symhip_problem_create.argtypes = [C.POINTER(ProblemDesc), C.POINTER(vp)]
        L.symhip_problem_destroy.argtypes = [vp]
        L.symhip_problem_dims.argtypes = [vp, I64P]
        L.symhip_problem_hessian_structure.argtypes = [vp, I32P, I32P]
        L.symhip_problem_col_starts.argtypes = [vp, I32P]
        L.symhip_problem_col_starts.restype = C.c_int64
        L.symhip_problem_state_keys.argtypes = [vp, I32P, I32P]
        L.symhip_linearize.argtypes = [vp, F64P, F64P, F64P, F64P]
        for name, args, res in _EXTRA_SIGNATURES:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _hip_lib = L
    return _hip_lib


class BatchTables(C.Structure):
    """symhip_batch_tables: device pointers of a packed batch (symhip_batch_solve_noalloc)."""
    _fields_ = [
        ("n_problems", C.c_int32),
        ("total_poses", C.c_int64),
        ("storage_size", C.c_int64),
        ("pose_begin", C.c_void_p),
        ("pose_soff", C.c_void_p),
        ("match_ptr", C.c_void_p),
        ("match_args", C.c_void_p),
        ("odo_args", C.c_void_p),
        ("values", C.c_void_p),
    ]


BATCH_NOALLOC_MAX_POSES = 16384


def batch_pack(problems, params: OptimizerParams | None = None):
    """Host tables of a batch in the batched kernel's layout (symhip_batch_pack); no device
    memory is touched."""
    params = params or default_params()
    n = len(problems)
    descs = (ProblemDesc * n)(*[p.desc() for p in problems])
    sizes = np.zeros(4, np.int64)
    check(hip_lib().symhip_batch_pack_sizes(descs, n, C.byref(params), ptr(sizes, C.c_int64)))
    _, poses, matches, storage = (int(x) for x in sizes)
    out = dict(pose_begin=np.zeros(n + 1, np.int32), pose_soff=np.zeros(poses, np.int32),
               match_ptr=np.zeros(poses + 1, np.int32),
               match_args=np.zeros(max(1, 3 * matches), np.int32),
               odo_args=np.zeros(3 * poses, np.int32), values=np.zeros(storage))
    check(hip_lib().symhip_batch_pack(
        descs, n, C.byref(params), *[ptr(out[k], C.c_int32) for k in
                                      ("pose_begin", "pose_soff", "match_ptr", "match_args",
                                       "odo_args")], ptr(out["values"], C.c_double)))
    out.update(n_problems=n, total_poses=poses, n_matches=matches, storage_size=storage)
    return out


def batch_solve_noalloc(tables: BatchTables, params, precision, epsilon, values_out, stats_out,
                        iters_out=None, stream=None):
    """symhip_batch_solve_noalloc on device pointers (ints); launches without synchronizing."""
    check(hip_lib().symhip_batch_solve_noalloc(C.byref(tables), C.byref(params), precision,
                                               epsilon, values_out, stats_out, iters_out, stream))


class SolverOptions(C.Structure):
    _fields_ = [
        ("linear_solver", C.c_int32),
        ("ordering", C.c_int32),
        ("precision", C.c_int32),
        ("use_pcg", C.c_int32),
        ("pcg_max_iters", C.c_int32),
        ("pcg_rel_tol", C.c_double),
        ("pcg_preconditioner", C.c_int32),
        ("pcg_exact_below_reduction", C.c_double),
        ("pcg_forcing", C.c_double),
    ]


SOLVER_CHOLESKY, SOLVER_SCHUR = 0, 1
PHASE_NAMES = ["linearize", "damp", "point_elim", "rcs_assemble", "reduced_rhs", "factor",
               "solve", "back_sub", "retract", "decide", "lin_points", "lin_cams"]
NUM_PHASES = len(PHASE_NAMES)
ORDERING_NATURAL, ORDERING_MIN_DEGREE, ORDERING_NESTED_DISSECTION = 0, 1, 2
PRECISION_F64, PRECISION_F32_JAC, PRECISION_F32_UPDATES = 0, 1, 2
PRECOND_BLOCK_JACOBI, PRECOND_LEARNED = 0, 1
LEARNED_PRECOND_PARAMS = 21744

_vp = C.c_void_p
_EXTRA_SIGNATURES: list = [
    ("symhip_optimizer_create", [C.POINTER(ProblemDesc), C.POINTER(OptimizerParams),
                                 C.POINTER(SolverOptions), C.c_double, C.POINTER(_vp)], C.c_int),
    ("symhip_optimizer_destroy", [_vp], C.c_int),
    ("symhip_optimizer_optimize", [_vp, F64P, C.c_int32, C.POINTER(IterationStats), C.c_int32,
                                   C.POINTER(OptimizationStats)], C.c_int),
    ("symhip_optimizer_optimize_gnc", [_vp, F64P, C.c_int32, C.POINTER(GncParams), C.c_int32,
                                       C.POINTER(IterationStats), C.c_int32,
                                       C.POINTER(OptimizationStats)], C.c_int),
    ("symhip_optimizer_best_linearization", [_vp, F64P, F64P, F64P], C.c_int),
    ("symhip_optimizer_warnings", [_vp, C.c_char_p, C.c_int64], C.c_int64),
    ("symhip_optimizer_jacobian_sparsity", [_vp, I32P, I32P], C.c_int64),
    ("symhip_optimizer_best_jacobian", [_vp, F64P], C.c_int),
    ("symhip_optimizer_debug_jacobian", [_vp, C.c_int32, F64P], C.c_int),
    ("symhip_optimizer_update_params", [_vp, C.POINTER(OptimizerParams)], C.c_int),
    ("symhip_optimizer_set_values", [_vp, F64P], C.c_int),
    ("symhip_optimizer_optimize_resident", [_vp, C.c_int32, C.POINTER(IterationStats), C.c_int32,
                                            C.POINTER(OptimizationStats)], C.c_int),
    ("symhip_optimizer_get_values", [_vp, F64P], C.c_int),
    ("symhip_optimizer_solver_info", [_vp, I64P], C.c_int),
    ("symhip_optimizer_linear_solver_stats", [_vp, I64P], C.c_int),
    ("symhip_optimizer_log_systems", [_vp, C.c_int32], C.c_int),
    ("symhip_optimizer_logged_systems", [_vp], C.c_int32),
    ("symhip_optimizer_rcs_pattern", [_vp, I32P, I32P], C.c_int32),
    ("symhip_optimizer_logged_system", [_vp, C.c_int32, F64P, F64P], C.c_int),
    ("symhip_optimizer_preconditioner_params", [_vp, F32P, F32P], C.c_int),
    ("symhip_optimizer_preconditioner_loss_grad", [_vp, C.c_int32, F64P, C.c_int32, F32P, F32P],
     C.c_int),
    ("symhip_optimizer_train_preconditioner", [_vp, C.c_int32, C.c_float, C.c_int32, C.c_uint64,
                                               C.c_int32, F32P], C.c_int),
    ("symhip_optimizer_eval_preconditioner", [_vp, C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                              I32P, F64P], C.c_int),
    ("symhip_batch_create", [C.POINTER(ProblemDesc), C.c_int32, C.POINTER(OptimizerParams),
                             C.c_int32, C.c_double, C.POINTER(_vp)], C.c_int),
    ("symhip_batch_destroy", [_vp], C.c_int),
    ("symhip_batch_storage_size", [_vp], C.c_int64),
    ("symhip_batch_solve", [_vp, F64P, C.POINTER(OptimizationStats), C.POINTER(IterationStats),
                            F64P, F64P], C.c_int),
    ("symhip_batch_tune", [_vp, C.c_int32, F64P, I32P], C.c_int),
    ("symhip_batch_set_config", [_vp, C.c_int32], C.c_int),
    ("symhip_batch_pack_sizes", [C.POINTER(ProblemDesc), C.c_int32, C.POINTER(OptimizerParams),
                                 I64P], C.c_int),
    ("symhip_batch_pack", [C.POINTER(ProblemDesc), C.c_int32, C.POINTER(OptimizerParams), I32P,
                           I32P, I32P, I32P, I32P, F64P], C.c_int),
    ("symhip_batch_solve_noalloc", [C.POINTER(BatchTables), C.POINTER(OptimizerParams), C.c_int32,
                                    C.c_double, _vp, _vp, _vp, _vp], C.c_int),
    ("symhip_optimizer_iteration_times", [_vp, F64P, C.c_int32], C.c_int32),
    ("symhip_optimizer_enable_timing", [_vp, C.c_int32], C.c_int),
    ("symhip_optimizer_phase_times", [_vp, F64P, I64P], C.c_int),
    ("symhip_cholesky_create", [C.c_int32, I32P, I32P, I32P, C.c_int32, C.c_int32, I32P,
                                C.POINTER(_vp)], C.c_int),
    ("symhip_cholesky_destroy", [_vp], C.c_int),
    ("symhip_cholesky_factorize", [_vp, F64P], C.c_int),
    ("symhip_cholesky_solve", [_vp, F64P, F64P, C.c_int32], C.c_int),
    ("symhip_cholesky_permutation", [_vp, I32P], C.c_int),
    ("symhip_cholesky_etree", [_vp, I32P, I32P], C.c_int),
    ("symhip_cholesky_d", [_vp, F64P], C.c_int),
    ("symhip_cholesky_info", [_vp, I64P], C.c_int),
    ("symhip_cholesky_factor", [_vp, I32P, I32P, F64P], C.c_int64),
    ("symhip_optimizer_debug_record", [_vp, C.c_int32, F64P, F64P, F64P], C.c_int),
    ("symhip_optimizer_factor_sparsity", [_vp, I32P, I32P, I32P], C.c_int64),
    ("symhip_factor_info", [C.c_int32, I32P], C.c_int),
    ("symhip_retract", [C.c_int32, I32P, I32P, I32P, F64P, F64P, C.c_double], C.c_int),
]


def factor_linearize_host(ftype, args, precision=PRECISION_F64):
    """symhip_factor_linearize_host: one registered factor body on the host; args is the list of
    argument storage arrays. Returns (residual (D,), Jacobian (D, T))."""
    info = np.zeros(4, np.int32)
    check(hip_lib().symhip_factor_info(int(ftype), ptr(info, C.c_int32)))
    arrs = [np.ascontiguousarray(a, np.float64) for a in args]
    assert len(arrs) == info[0], "argument count"
    ptrs = (F64P * len(arrs))(*[a.ctypes.data_as(F64P) for a in arrs])
    r = np.zeros(info[2])
    J = np.zeros(info[2] * info[3])
    check(hip_lib().symhip_factor_linearize_host(int(ftype), ptrs, precision,
                                                 ptr(r, C.c_double), ptr(J, C.c_double)))
    return r, J.reshape(info[2], info[3])


def check(rc):
    if rc != 0:
        raise RuntimeError(hip_lib().symhip_last_error().decode())


class HipProblem:
    """Problem handle of libsymhip (structure on the host, device state on first use)."""

    def __init__(self, prob: Problem):
        self.prob = prob
        self._desc = prob.desc()
        h = C.c_void_p()
        check(hip_lib().symhip_problem_create(C.byref(self._desc), C.byref(h)))
        self.h = h
        dims = np.zeros(5, np.int64)
        check(hip_lib().symhip_problem_dims(self.h, ptr(dims, C.c_int64)))
        self.N, self.M, self.nnz, self.storage, self.n_state = (int(x) for x in dims)

    def __del__(self):
        if getattr(self, "h", None):
            hip_lib().symhip_problem_destroy(self.h)
            self.h = None

    def hessian_structure(self):
        cp = np.zeros(self.N + 1, np.int32)
        ri = np.zeros(self.nnz, np.int32)
        check(hip_lib().symhip_problem_hessian_structure(self.h, ptr(cp, C.c_int32),
                                                         ptr(ri, C.c_int32)))
        return cp, ri

    def col_starts(self):
        n = hip_lib().symhip_problem_col_starts(self.h, None)
        out = np.zeros(n, np.int32)
        hip_lib().symhip_problem_col_starts(self.h, ptr(out, C.c_int32))
        return out

    def state_keys(self):
        k = np.zeros(self.n_state, np.int32)
        o = np.zeros(self.n_state, np.int32)
        check(hip_lib().symhip_problem_state_keys(self.h, ptr(k, C.c_int32), ptr(o, C.c_int32)))
        return k, o

    def linearize(self, values=None):
        v = self.prob.values if values is None else np.ascontiguousarray(values, np.float64)
        r, h, b = np.zeros(self.M), np.zeros(self.nnz), np.zeros(self.N)
        check(hip_lib().symhip_linearize(self.h, ptr(v, C.c_double), ptr(r, C.c_double),
                                         ptr(h, C.c_double), ptr(b, C.c_double)))
        return dict(residual=r, hessian_values=h, rhs=b)


def solver_options(linear_solver=SOLVER_CHOLESKY, ordering=ORDERING_MIN_DEGREE,
                   precision=PRECISION_F64, use_pcg=0, pcg_max_iters=100, pcg_rel_tol=1e-6,
                   pcg_preconditioner=0, pcg_exact_below_reduction=0.0, pcg_forcing=0.0):
    return SolverOptions(linear_solver, ordering, precision, use_pcg, pcg_max_iters, pcg_rel_tol,
                         pcg_preconditioner, pcg_exact_below_reduction, pcg_forcing)


class HipOptimizer:
    """sym::Optimizer<double> over libsymhip (device-resident LM)."""

    def __init__(self, prob: Problem, params: OptimizerParams | None = None, options=None,
                 epsilon=None):
        self.prob = prob
        self.params = params or default_params()
        self.options = options or solver_options()
        eps = prob.info.get("epsilon", 10 * np.finfo(np.float64).eps) if epsilon is None else epsilon
        self._desc = prob.desc()
        h = C.c_void_p()
        check(hip_lib().symhip_optimizer_create(C.byref(self._desc), C.byref(self.params),
                                                C.byref(self.options), eps, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            hip_lib().symhip_optimizer_destroy(self.h)
            self.h = None

    def optimize(self, values=None, num_iterations=-1):
        vals = (self.prob.values if values is None else np.asarray(values, np.float64)).copy()
        maxit = 4096
        iters = (IterationStats * maxit)()
        st = OptimizationStats()
        check(hip_lib().symhip_optimizer_optimize(self.h, ptr(vals, C.c_double), num_iterations,
                                                  iters, maxit, C.byref(st)))
        return dict(values=vals, status=st.status, failure_reason=st.failure_reason,
                    best_index=st.best_index,
                    iterations=iteration_records(iters, st.num_iterations))

    def optimize_gnc(self, gnc, mu_key, values=None, num_iterations=-1):
        """GncOptimizer::Optimize (opt/gnc_optimizer.h:49-118); mu_key indexes the problem's
        keys."""
        vals = (self.prob.values if values is None else np.asarray(values, np.float64)).copy()
        maxit = 4096
        iters = (IterationStats * maxit)()
        st = OptimizationStats()
        check(hip_lib().symhip_optimizer_optimize_gnc(self.h, ptr(vals, C.c_double),
                                                      num_iterations, C.byref(gnc), mu_key,
                                                      iters, maxit, C.byref(st)))
        return dict(values=vals, status=st.status, failure_reason=st.failure_reason,
                    best_index=st.best_index,
                    iterations=iteration_records(iters, st.num_iterations))

    def jacobian_sparsity(self, N):
        """include_jacobians: the Jacobian's CSC pattern (colptr over the N columns, rowind)."""
        L = hip_lib()
        nnz = L.symhip_optimizer_jacobian_sparsity(self.h, None, None)
        check(nnz if nnz < 0 else 0)
        cp = np.zeros(N + 1, np.int32)
        ri = np.zeros(nnz, np.int32)
        check(min(0, L.symhip_optimizer_jacobian_sparsity(self.h, ptr(cp, C.c_int32),
                                                          ptr(ri, C.c_int32))))
        return cp, ri

    def best_jacobian(self, nnz):
        v = np.zeros(nnz)
        check(hip_lib().symhip_optimizer_best_jacobian(self.h, ptr(v, C.c_double)))
        return v

    def debug_jacobian(self, index, nnz):
        v = np.zeros(nnz)
        check(hip_lib().symhip_optimizer_debug_jacobian(self.h, index, ptr(v, C.c_double)))
        return v

    def warnings(self):
        """The last optimize call's warnings (symhip_optimizer_warnings), one string each."""
        n = hip_lib().symhip_optimizer_warnings(self.h, None, 0)
        check(n if n < 0 else 0)
        buf = C.create_string_buffer(int(n) + 1)
        check(min(0, hip_lib().symhip_optimizer_warnings(self.h, buf, n + 1)))
        return [w for w in buf.value.decode().split("\x1e") if w]

    def set_values(self, values):
        v = np.ascontiguousarray(values, np.float64)
        check(hip_lib().symhip_optimizer_set_values(self.h, ptr(v, C.c_double)))

    def optimize_resident(self, num_iterations=-1):
        maxit = 4096
        iters = (IterationStats * maxit)()
        st = OptimizationStats()
        check(hip_lib().symhip_optimizer_optimize_resident(self.h, num_iterations, iters, maxit,
                                                           C.byref(st)))
        return dict(status=st.status, failure_reason=st.failure_reason, best_index=st.best_index,
                    iterations=iteration_records(iters, st.num_iterations))

    def get_values(self):
        v = np.zeros(self.prob.values.size)
        check(hip_lib().symhip_optimizer_get_values(self.h, ptr(v, C.c_double)))
        return v

    def best_linearization(self, M, nnz, N):
        r, h, b = np.zeros(M), np.zeros(nnz), np.zeros(N)
        check(hip_lib().symhip_optimizer_best_linearization(
            self.h, ptr(r, C.c_double), ptr(h, C.c_double), ptr(b, C.c_double)))
        return dict(residual=r, hessian_values=h, rhs=b)

    def iteration_times(self):
        out = np.zeros(4096)
        n = hip_lib().symhip_optimizer_iteration_times(self.h, ptr(out, C.c_double), 4096)
        return out[:n]

    def enable_timing(self, on=True):
        check(hip_lib().symhip_optimizer_enable_timing(self.h, int(on)))

    def phase_times(self):
        ms = np.zeros(NUM_PHASES)
        cnt = np.zeros(NUM_PHASES, np.int64)
        check(hip_lib().symhip_optimizer_phase_times(self.h, ptr(ms, C.c_double),
                                                     ptr(cnt, C.c_int64)))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(PHASE_NAMES)}

    def solver_info(self):
        out = np.zeros(10, np.int64)
        check(hip_lib().symhip_optimizer_solver_info(self.h, ptr(out, C.c_int64)))
        return dict(n=int(out[0]), supernodes=int(out[1]), panel=int(out[2]),
                    levels=int(out[3]), flops=int(out[4]), blocks=int(out[5]),
                    pairs=int(out[6]), obs=int(out[7]), diag_panel_flops=int(out[8]),
                    update_flops=int(out[9]))

    # ---- learned PCG preconditioner (Schur + use_pcg) ----
    def log_systems(self, max_systems):
        check(hip_lib().symhip_optimizer_log_systems(self.h, max_systems))

    def logged_systems(self):
        return int(hip_lib().symhip_optimizer_logged_systems(self.h))

    def rcs_pattern(self):
        n = hip_lib().symhip_optimizer_rcs_pattern(self.h, None, None)
        if n < 0:
            check(-1)
        row_ptr = np.zeros(n + 1, np.int32)
        hip_lib().symhip_optimizer_rcs_pattern(self.h, ptr(row_ptr, C.c_int32), None)
        col = np.zeros(int(row_ptr[-1]), np.int32)
        hip_lib().symhip_optimizer_rcs_pattern(self.h, ptr(row_ptr, C.c_int32),
                                               ptr(col, C.c_int32))
        return row_ptr, col

    def logged_system(self, idx):
        row_ptr, col = self.rcs_pattern()
        blocks = np.zeros(len(col) * 81)
        rhs = np.zeros((len(row_ptr) - 1) * 9)
        check(hip_lib().symhip_optimizer_logged_system(self.h, idx, ptr(blocks, C.c_double),
                                                       ptr(rhs, C.c_double)))
        return row_ptr, col, blocks, rhs

    def preconditioner_params(self, set_params=None):
        out = np.zeros(LEARNED_PRECOND_PARAMS, np.float32)
        src = None if set_params is None else np.ascontiguousarray(set_params, np.float32)
        check(hip_lib().symhip_optimizer_preconditioner_params(
            self.h, ptr(out, C.c_float), None if src is None else ptr(src, C.c_float)))
        return out

    def preconditioner_loss_grad(self, idx, X):
        X = np.asfortranarray(np.asarray(X, np.float64))  # probe p contiguous
        P = X.shape[1]
        loss = np.zeros(1, np.float32)
        grad = np.zeros(LEARNED_PRECOND_PARAMS, np.float32)
        check(hip_lib().symhip_optimizer_preconditioner_loss_grad(
            self.h, idx, X.ctypes.data_as(F64P), P, ptr(loss, C.c_float), ptr(grad, C.c_float)))
        return float(loss[0]), grad

    def train_preconditioner(self, steps, lr=1e-3, probes=4, seed=0, n_systems=0):
        losses = np.zeros(steps, np.float32)
        check(hip_lib().symhip_optimizer_train_preconditioner(self.h, steps, lr, probes, seed,
                                                              n_systems, ptr(losses, C.c_float)))
        return losses

    def eval_preconditioner(self, idx, rel_tol=1e-6, max_iters=1000, learned=False):
        it = np.zeros(1, np.int32)
        err = np.zeros(1)
        check(hip_lib().symhip_optimizer_eval_preconditioner(
            self.h, idx, rel_tol, max_iters, int(learned), ptr(it, C.c_int32),
            ptr(err, C.c_double)))
        return int(it[0]), float(err[0])

    def refinement_residuals(self, track=-1, max_count=4096):
        """Mixed-precision sparse Cholesky: track 1/0 turns the per-solve record of the final
        relative residual |b - A x| / |b| on/off (and clears it); returns the recorded values."""
        out = np.zeros(max_count)
        n = hip_lib().symhip_optimizer_refinement_residuals(self.h, track, ptr(out, C.c_double),
                                                            max_count)
        if n < 0:
            check(-1)
        return out[:min(n, max_count)]

    def linear_solver_stats(self):
        out = np.zeros(4, np.int64)
        check(hip_lib().symhip_optimizer_linear_solver_stats(self.h, ptr(out, C.c_int64)))
        return dict(solves=int(out[0]), iterations=int(out[1]), fallbacks=int(out[2]),
                    max_iterations=int(out[3]))


class HipBatch:
    """Many small localization problems solved by one persistent launch (one wavefront each)."""

    def __init__(self, problems, params: OptimizerParams | None = None, precision=PRECISION_F64,
                 epsilon=None):
        self.problems = list(problems)
        self.params = params or default_params()
        eps = (self.problems[0].info.get("epsilon", 10 * np.finfo(np.float64).eps)
               if epsilon is None else epsilon)
        n = len(self.problems)
        self._descs = (ProblemDesc * n)(*[p.desc() for p in self.problems])
        h = C.c_void_p()
        check(hip_lib().symhip_batch_create(self._descs, n, C.byref(self.params), precision, eps,
                                            C.byref(h)))
        self.h = h
        self.sizes = [len(p.values) for p in self.problems]

    def __del__(self):
        if getattr(self, "h", None):
            hip_lib().symhip_batch_destroy(self.h)
            self.h = None

    BATCH_CONFIGS = ("1 wave/SIMD, given order", "2 waves/SIMD, given order",
                     "1 wave/SIMD, most poses first", "2 waves/SIMD, most poses first")

    def tune(self, reps=2):
        """The batch autotuner (symhip_batch_tune): times each launch configuration over `reps`
        solves, keeps the fastest; returns (chosen index, per-configuration ms)."""
        ms = np.zeros(4)
        chosen = np.zeros(1, np.int32)
        check(hip_lib().symhip_batch_tune(self.h, reps, ptr(ms, C.c_double),
                                          ptr(chosen, C.c_int32)))
        return int(chosen[0]), ms

    def set_config(self, config):
        check(hip_lib().symhip_batch_set_config(self.h, config))

    def solve(self, records=True):
        n = len(self.problems)
        vals = np.zeros(sum(self.sizes))
        stats = (OptimizationStats * n)()
        stride = self.params.iterations + 1
        iters = (IterationStats * (n * stride))() if records else None
        ms = np.zeros(n)
        total = np.zeros(1)
        check(hip_lib().symhip_batch_solve(self.h, ptr(vals, C.c_double), stats, iters,
                                           ptr(ms, C.c_double), ptr(total, C.c_double)))
        out, off = [], 0
        for i in range(n):
            st = stats[i]
            rec = None
            if records:
                sub = (IterationStats * stride).from_address(C.addressof(iters) +
                                                             i * stride * C.sizeof(IterationStats))
                rec = iteration_records(sub, st.num_iterations)
            out.append(dict(values=vals[off:off + self.sizes[i]].copy(), status=st.status,
                            failure_reason=st.failure_reason, best_index=st.best_index,
                            iterations=rec))
            off += self.sizes[i]
        return dict(problems=out, solve_ms=ms, total_ms=float(total[0]))


def gen_localization_batch(n_problems=4096, seed=42, min_poses=5, max_poses=50,
                           min_landmarks=5, max_landmarks=40, visible_fraction=0.3):
    """Independent robot-3D-localization problems with uniform pose / landmark counts."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_problems):
        npose = int(rng.integers(min_poses, max_poses + 1))
        nlm = int(rng.integers(min_landmarks, max_landmarks + 1))
        out.append(gen_localization(npose, nlm, visible_fraction, seed=int(rng.integers(1 << 31))))
    return out


class HipCholesky:
    """SparseCholeskySolver replacement (supernodal LDL^T on the device)."""

    def __init__(self, n, colptr, rowind, node_ptr=None, ordering=ORDERING_MIN_DEGREE,
                 given_order=None):
        self.n = n
        self.colptr = np.ascontiguousarray(colptr, np.int32)
        self.rowind = np.ascontiguousarray(rowind, np.int32)
        npv = None if node_ptr is None else np.ascontiguousarray(node_ptr, np.int32)
        gv = None if given_order is None else np.ascontiguousarray(given_order, np.int32)
        h = C.c_void_p()
        check(hip_lib().symhip_cholesky_create(
            n, ptr(self.colptr, C.c_int32), ptr(self.rowind, C.c_int32),
            None if npv is None else ptr(npv, C.c_int32), 0 if npv is None else npv.size - 1,
            ordering, None if gv is None else ptr(gv, C.c_int32), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            hip_lib().symhip_cholesky_destroy(self.h)
            self.h = None

    def set_f32_updates(self, on=True):
        """Mixed-precision factorization (trailing updates on the FP32 matrix cores)."""
        check(hip_lib().symhip_cholesky_set_f32_updates(self.h, int(on)))

    def inv_permutation(self):
        out = np.zeros(self.n, np.int32)
        check(hip_lib().symhip_cholesky_permutation(self.h, ptr(out, C.c_int32)))
        return out

    def etree(self):
        p = np.zeros(self.n, np.int32)
        c = np.zeros(self.n, np.int32)
        check(hip_lib().symhip_cholesky_etree(self.h, ptr(p, C.c_int32), ptr(c, C.c_int32)))
        return p, c

    def factor_l(self):
        """L() after a factorization: (colptr, rowind, values) of the strictly lower unit factor
        of the permuted matrix on the reference's pattern."""
        nnz = hip_lib().symhip_cholesky_factor(self.h, None, None, None)
        if nnz < 0:
            check(int(nnz))
        cp = np.zeros(self.n + 1, np.int32)
        ri = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        got = hip_lib().symhip_cholesky_factor(self.h, ptr(cp, C.c_int32), ptr(ri, C.c_int32),
                                                ptr(v, C.c_double))
        if got < 0:
            check(int(got))
        return cp, ri, v

    def info(self):
        out = np.zeros(5, np.int64)
        check(hip_lib().symhip_cholesky_info(self.h, ptr(out, C.c_int64)))
        return dict(n=int(out[0]), supernodes=int(out[1]), panel=int(out[2]),
                    levels=int(out[3]), flops=int(out[4]))

    def factorize(self, values):
        v = np.ascontiguousarray(values, np.float64)
        check(hip_lib().symhip_cholesky_factorize(self.h, ptr(v, C.c_double)))

    def solve(self, b):
        b = np.asarray(b, np.float64)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        bb = np.ascontiguousarray(b.T if b.ndim == 2 else b)
        x = np.zeros_like(bb)
        check(hip_lib().symhip_cholesky_solve(self.h, ptr(bb, C.c_double), ptr(x, C.c_double),
                                              nrhs))
        return x.T if b.ndim == 2 else x

    def d(self):
        out = np.zeros(self.n)
        check(hip_lib().symhip_cholesky_d(self.h, ptr(out, C.c_double)))
        return out


def compute_full_covariance(problem: "HipProblem", hessian_values):
    """Optimizer::ComputeFullCovariance (opt/optimizer.tcc:202-206) through
    LevenbergMarquardtSolver::ComputeCovariance (opt/levenberg_marquardt_solver.tcc:345-356): the
    undamped lower Hessian of a linearization (values in the problem's CSC order) is factored by
    the device LDL^T and solved against the identity; the result is the dense N x N inverse.
    Raises, like the reference's assert, when the factorization fails."""
    cp, ri = problem.hessian_structure()
    chol = HipCholesky(problem.N, cp, ri)
    chol.factorize(hessian_values)
    cov = chol.solve(np.eye(problem.N))
    if not np.isfinite(cov).all():
        raise RuntimeError("Internal Error: damped hessian factorization failed")
    return cov


def compute_all_covariances(problem: "HipProblem", hessian_values):
    """Optimizer::ComputeAllCovariances (opt/optimizer.tcc:113-121): the full covariance split by
    key as SplitCovariancesByKey does (opt/internal/covariance_utils.h:172-192), the diagonal
    block at each optimized key's tangent offset. Returns {key index in the problem: block}."""
    cov = compute_full_covariance(problem, hessian_values)
    keys, offsets = problem.state_keys()
    ends = np.append(offsets[1:], problem.N)
    return {int(k): cov[o:e, o:e].copy() for k, o, e in zip(keys, offsets, ends)}
