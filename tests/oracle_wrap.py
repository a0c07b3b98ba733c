"""ctypes wrapper of the CPU FP64 oracle (oracle/oracle.h). Test infrastructure only: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the package."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from symforce_amd import _capi as capi

REPO = Path(__file__).resolve().parent.parent
ORACLE_LIB = REPO / "oracle" / "_build" / "liboracle.so"
# the same sources built with -ffast-math (the reference's recommended benchmark flags,
# symforce/benchmarks/README.md:15); used only to time the CPU baseline
ORACLE_FASTMATH_LIB = REPO / "oracle" / "_build" / "liboracle_fastmath.so"
CHOLESKY, SCHUR = 0, 1

_libs = {}


def lib(fast_math=False):
    path = ORACLE_FASTMATH_LIB if fast_math else ORACLE_LIB
    if path not in _libs:
        if not path.exists():
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.oracle_problem_create.argtypes = [C.POINTER(capi.ProblemDesc), C.c_int]
        L.oracle_problem_create.restype = vp
        L.oracle_problem_destroy.argtypes = [vp]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_dims.argtypes = [vp, capi.I64P]
        L.oracle_hessian_structure.argtypes = [vp, capi.I32P, capi.I32P]
        L.oracle_jacobian_structure.argtypes = [vp, capi.I32P, capi.I32P]
        L.oracle_storage_offsets.argtypes = [vp, capi.I32P]
        L.oracle_state_keys.argtypes = [vp, capi.I32P, capi.I32P]
        L.oracle_hessian_col_starts.argtypes = [vp, capi.I32P]
        L.oracle_hessian_col_starts.restype = C.c_int64
        L.oracle_linearize.argtypes = [vp, capi.F64P, capi.F64P, capi.F64P, capi.F64P, capi.F64P]
        L.oracle_cholesky_create.argtypes = [C.c_int32, capi.I32P, capi.I32P, capi.I32P]
        L.oracle_cholesky_create.restype = vp
        L.oracle_cholesky_destroy.argtypes = [vp]
        L.oracle_cholesky_factorize.argtypes = [vp, capi.F64P]
        L.oracle_cholesky_solve.argtypes = [vp, capi.F64P, capi.F64P, C.c_int32]
        L.oracle_cholesky_nnz.argtypes = [vp]
        L.oracle_cholesky_nnz.restype = C.c_int64
        L.oracle_cholesky_factor.argtypes = [vp, capi.I32P, capi.I32P, capi.F64P, capi.F64P,
                                             capi.I32P]
        L.oracle_schur_create.argtypes = [C.c_int32, capi.I32P, capi.I32P, C.c_int32, capi.I32P]
        L.oracle_schur_create.restype = vp
        L.oracle_schur_destroy.argtypes = [vp]
        L.oracle_schur_factorize.argtypes = [vp, capi.F64P]
        L.oracle_schur_solve.argtypes = [vp, capi.F64P, capi.F64P]
        L.oracle_schur_dense_s.argtypes = [vp, capi.F64P]
        L.oracle_optimize.argtypes = [vp, C.POINTER(capi.OptimizerParams), C.c_int, capi.I32P,
                                      C.c_int32, C.c_double, C.c_int32, capi.F64P,
                                      C.POINTER(capi.IterationStats), C.c_int32,
                                      C.POINTER(capi.OptimizationStats), capi.F64P]
        L.oracle_optimize_timed.argtypes = [vp, C.POINTER(capi.OptimizerParams), C.c_int,
                                            capi.I32P, C.c_int32, C.c_double, C.c_int32,
                                            capi.F64P, C.POINTER(capi.IterationStats), C.c_int32,
                                            C.POINTER(capi.OptimizationStats), capi.F64P]
        L.oracle_optimize_gnc.argtypes = [vp, C.POINTER(capi.OptimizerParams), C.c_int,
                                          C.c_double, C.c_int32, C.POINTER(capi.GncParams),
                                          C.c_int32, capi.F64P, C.POINTER(capi.IterationStats),
                                          C.c_int32, C.POINTER(capi.OptimizationStats)]
        L.oracle_linearize_factor.argtypes = [C.c_int32, C.POINTER(capi.F64P), capi.F64P,
                                              capi.F64P, capi.F64P, capi.F64P]
        _libs[path] = L
    return _libs[path]


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


class OracleProblem:
    def __init__(self, prob: capi.Problem, include_jacobians=False, fast_math=False):
        self.prob = prob
        self._desc = prob.desc()
        self.L = lib(fast_math)
        self.h = self.L.oracle_problem_create(C.byref(self._desc), int(include_jacobians))
        if not self.h:
            raise RuntimeError(self.L.oracle_last_error().decode())
        dims = np.zeros(4, np.int64)
        self.L.oracle_dims(self.h, _p(dims, C.c_int64))
        self.N, self.M, self.nnz, self.jnnz = (int(x) for x in dims)
        self.include_jacobians = include_jacobians

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_problem_destroy(self.h)
            self.h = None

    def hessian_structure(self):
        cp = np.zeros(self.N + 1, np.int32)
        ri = np.zeros(self.nnz, np.int32)
        self.L.oracle_hessian_structure(self.h, _p(cp, C.c_int32), _p(ri, C.c_int32))
        return cp, ri

    def jacobian_structure(self):
        cp = np.zeros(self.N + 1, np.int32)
        ri = np.zeros(self.jnnz, np.int32)
        self.L.oracle_jacobian_structure(self.h, _p(cp, C.c_int32), _p(ri, C.c_int32))
        return cp, ri

    def state_keys(self):
        k = np.zeros(self.prob.n_keys, np.int32)
        o = np.zeros(self.prob.n_keys, np.int32)
        n = self.L.oracle_state_keys(self.h, _p(k, C.c_int32), _p(o, C.c_int32))
        return k[:n], o[:n]

    def hessian_col_starts(self):
        n = self.L.oracle_hessian_col_starts(self.h, None)
        out = np.zeros(n, np.int32)
        self.L.oracle_hessian_col_starts(self.h, _p(out, C.c_int32))
        return out

    def linearize(self, values=None):
        v = self.prob.values if values is None else np.ascontiguousarray(values, np.float64)
        r = np.zeros(self.M)
        h = np.zeros(self.nnz)
        b = np.zeros(self.N)
        j = np.zeros(self.jnnz) if self.include_jacobians else None
        rc = self.L.oracle_linearize(self.h, _p(v, C.c_double), _p(r, C.c_double),
                                    _p(h, C.c_double), _p(b, C.c_double), _p(j, C.c_double))
        if rc != 0:
            raise RuntimeError(self.L.oracle_last_error().decode())
        return dict(residual=r, hessian_values=h, rhs=b, jacobian_values=j)

    def optimize(self, params=None, solver=CHOLESKY, ordering=None, c_dim=0, epsilon=None,
                 num_iterations=-1, values=None, record_updates=False):
        params = params or capi.default_params()
        eps = self.prob.info.get("epsilon", 2.220446049250313e-15) if epsilon is None else epsilon
        vals = (self.prob.values if values is None else np.asarray(values, np.float64)).copy()
        maxit = 4096
        iters = (capi.IterationStats * maxit)()
        st = capi.OptimizationStats()
        ordv = None if ordering is None else np.ascontiguousarray(ordering, np.int32)
        nmax = (params.iterations if num_iterations < 0 else num_iterations) + 1
        upd = np.zeros(nmax * self.N) if record_updates else None
        rc = self.L.oracle_optimize(self.h, C.byref(params), solver, _p(ordv, C.c_int32), c_dim,
                                   eps, num_iterations, _p(vals, C.c_double), iters, maxit,
                                   C.byref(st), _p(upd, C.c_double))
        if rc != 0:
            raise RuntimeError(self.L.oracle_last_error().decode())
        out = dict(values=vals, status=st.status, failure_reason=st.failure_reason,
                   best_index=st.best_index,
                   iterations=capi.iteration_records(iters, st.num_iterations))
        if record_updates:
            out["updates"] = upd[: (st.num_iterations - 1) * self.N].reshape(-1, self.N)
        return out


    def optimize_gnc(self, gnc, mu_key, params=None, solver=CHOLESKY, epsilon=None,
                     num_iterations=-1, values=None):
        params = params or capi.default_params()
        eps = self.prob.info.get("epsilon", 2.220446049250313e-15) if epsilon is None else epsilon
        vals = (self.prob.values if values is None else np.asarray(values, np.float64)).copy()
        maxit = 4096
        iters = (capi.IterationStats * maxit)()
        st = capi.OptimizationStats()
        rc = self.L.oracle_optimize_gnc(self.h, C.byref(params), solver, eps, num_iterations,
                                       C.byref(gnc), mu_key, _p(vals, C.c_double), iters, maxit,
                                       C.byref(st))
        if rc != 0:
            raise RuntimeError(self.L.oracle_last_error().decode())
        return dict(values=vals, status=st.status, failure_reason=st.failure_reason,
                    best_index=st.best_index,
                    iterations=capi.iteration_records(iters, st.num_iterations))


def time_iterations(op: "OracleProblem", params, num_iterations, solver=CHOLESKY, c_dim=0,
                    ordering=None):
    """Runs num_iterations LM iterations and returns the seconds spent in each Iterate()
    (ordering: elimination order of the factorized matrix, H or the Schur complement S)."""
    vals = op.prob.values.copy()
    iters = (capi.IterationStats * (num_iterations + 2))()
    st = capi.OptimizationStats()
    secs = np.zeros(num_iterations)
    eps = op.prob.info.get("epsilon", 2.220446049250313e-15)
    ordv = None if ordering is None else np.ascontiguousarray(ordering, np.int32)
    rc = op.L.oracle_optimize_timed(op.h, C.byref(params), solver, _p(ordv, C.c_int32), c_dim,
                                    eps, num_iterations, _p(vals, C.c_double), iters,
                                    num_iterations + 2, C.byref(st), _p(secs, C.c_double))
    if rc != 0:
        raise RuntimeError(op.L.oracle_last_error().decode())
    return secs, capi.iteration_records(iters, st.num_iterations)


class OracleCholesky:
    def __init__(self, n, colptr, rowind, ordering=None):
        self.n = n
        self.colptr = np.ascontiguousarray(colptr, np.int32)
        self.rowind = np.ascontiguousarray(rowind, np.int32)
        o = None if ordering is None else np.ascontiguousarray(ordering, np.int32)
        self.h = lib().oracle_cholesky_create(n, _p(self.colptr, C.c_int32),
                                              _p(self.rowind, C.c_int32), _p(o, C.c_int32))
        if not self.h:
            raise RuntimeError(lib().oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_cholesky_destroy(self.h)

    def factorize(self, values):
        v = np.ascontiguousarray(values, np.float64)
        lib().oracle_cholesky_factorize(self.h, _p(v, C.c_double))

    def solve(self, b):
        b = np.ascontiguousarray(b, np.float64)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        bb = np.ascontiguousarray(b.T if b.ndim == 2 else b)
        x = np.zeros_like(bb)
        lib().oracle_cholesky_solve(self.h, _p(bb, C.c_double), _p(x, C.c_double), nrhs)
        return x.T if b.ndim == 2 else x

    def factor(self):
        nnz = lib().oracle_cholesky_nnz(self.h)
        cp = np.zeros(self.n + 1, np.int32)
        ri = np.zeros(nnz, np.int32)
        lv = np.zeros(nnz)
        d = np.zeros(self.n)
        par = np.zeros(self.n, np.int32)
        lib().oracle_cholesky_factor(self.h, _p(cp, C.c_int32), _p(ri, C.c_int32),
                                     _p(lv, C.c_double), _p(d, C.c_double), _p(par, C.c_int32))
        return dict(colptr=cp, rowind=ri, values=lv, d=d, parent=par)


class OracleSchur:
    def __init__(self, n, colptr, rowind, c_dim, s_ordering=None):
        self.n, self.c_dim = n, c_dim
        self.colptr = np.ascontiguousarray(colptr, np.int32)
        self.rowind = np.ascontiguousarray(rowind, np.int32)
        o = None if s_ordering is None else np.ascontiguousarray(s_ordering, np.int32)
        self.h = lib().oracle_schur_create(n, _p(self.colptr, C.c_int32),
                                           _p(self.rowind, C.c_int32), c_dim, _p(o, C.c_int32))
        if not self.h:
            raise RuntimeError(lib().oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_schur_destroy(self.h)

    def factorize(self, values):
        v = np.ascontiguousarray(values, np.float64)
        if lib().oracle_schur_factorize(self.h, _p(v, C.c_double)) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def solve(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        lib().oracle_schur_solve(self.h, _p(b, C.c_double), _p(x, C.c_double))
        return x

    def dense_s(self):
        B = self.n - self.c_dim
        out = np.zeros(B * B)
        lib().oracle_schur_dense_s(self.h, _p(out, C.c_double))
        return out.reshape(B, B).T


def linearize_factor(ftype, args):
    """Single-factor linearization: args is a list of float64 storage arrays."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in args]
    ptrs = (capi.F64P * len(arrs))(*[_p(a, C.c_double) for a in arrs])
    lin = {}
    out = [np.zeros(64), np.zeros(64 * 64), np.zeros(64 * 64), np.zeros(64)]
    rc = lib().oracle_linearize_factor(ftype, ptrs, *[_p(o, C.c_double) for o in out])
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out, arrs


def csc_to_dense(n, colptr, rowind, vals, rows=None):
    rows = n if rows is None else rows
    A = np.zeros((rows, n))
    for c in range(n):
        for q in range(colptr[c], colptr[c + 1]):
            A[rowind[q], c] += vals[q]
    return A


def schur_elimination_order(prob):
    """Elimination order of the reduced camera system for the CPU oracle: nested dissection of
    the camera covisibility graph (the reference orders S with METIS, sparse_schur_solver.h:77),
    each camera's 6 pose and 3 intrinsics columns kept together. State order is all poses, then
    all intrinsics, then the points."""
    ft, fp, fa = (np.asarray(x) for x in (prob.ftypes, prob.fptr, prob.fargs))
    starts = fp[:-1][ft == capi.FACTOR_SNAVELY_HUBER]
    cam_key, pt_key = fa[starts], fa[starts + 2]
    ucam = np.unique(cam_key)
    cam = np.searchsorted(ucam, cam_key)
    o = np.lexsort((cam, pt_key))
    pk, ck = pt_key[o], cam[o]
    n = ucam.size
    rows, cols = [np.arange(n)], [np.arange(n)]
    for d in range(1, int(np.bincount(pk).max())):
        same = pk[d:] == pk[:-d]
        a, b = ck[:-d][same], ck[d:][same]
        rows.append(np.maximum(a, b))
        cols.append(np.minimum(a, b))
    key = np.unique(np.concatenate(cols).astype(np.int64) * n + np.concatenate(rows))
    c, r = key // n, key % n
    colptr = np.concatenate([[0], np.cumsum(np.bincount(c, minlength=n))]).astype(np.int32)
    order = capi.HipCholesky(n, colptr, r.astype(np.int32),
                             ordering=capi.ORDERING_NESTED_DISSECTION).inv_permutation()
    return np.concatenate([np.concatenate([6 * k + np.arange(6), 6 * n + 3 * k + np.arange(3)])
                           for k in order]).astype(np.int32)
