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
TYPE_VECTOR2, TYPE_VECTOR3, TYPE_VECTOR4, TYPE_VECTOR6 = 12, 13, 14, 16
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

LAMBDA_STATIC, LAMBDA_DYNAMIC = 1, 2
STATUS_INVALID, STATUS_SUCCESS, STATUS_HIT_ITERATION_LIMIT, STATUS_FAILED = 0, 1, 2, 3
FAILURE_INVALID, FAILURE_LAMBDA_OUT_OF_BOUNDS, FAILURE_INITIAL_ERROR_NOT_FINITE = 0, 1, 2

STORAGE_DIM = {1: 1, 3: 4, 5: 7, 12: 2, 13: 3, 14: 4, 16: 6, 37: 9, 64: 36}
TANGENT_DIM = {1: 1, 3: 3, 5: 6, 12: 2, 13: 3, 14: 4, 16: 6, 37: 9, 64: 36}

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
I8P = C.POINTER(C.c_int8)
F64P = C.POINTER(C.c_double)


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


def gen_ba_snavely(n_cams=1000, n_points=200000, seed=42):
    return _from_generated(problems_lib().symhip_gen_ba_snavely(n_cams, n_points, seed))


def gen_pose_graph(n_nodes=100000, n_loops=30000, seed=42):
    return _from_generated(problems_lib().symhip_gen_pose_graph(n_nodes, n_loops, seed))


def gen_localization(n_poses=20, n_landmarks=20, visible_fraction=0.3, seed=42):
    return _from_generated(
        problems_lib().symhip_gen_localization(n_poses, n_landmarks, visible_fraction, seed))
