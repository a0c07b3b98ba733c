"""symhip_batch_solve_noalloc (the allocation-free batched LM entry point) against
symhip_batch_solve on the same problems, through the C ABI: bit-identical Values, statistics and
iteration records, on two calls in a row (the second with a smaller batch, so stale state in the
shared scratch array would show), in both precisions; oversize batches are refused on the host.

Device buffers for the tables and outputs are the caller's: they are allocated here with the HIP
runtime that libsymhip.so itself links (libamdhip64.so.7), outside the entry point under test."""
import ctypes as C

import numpy as np
import pytest

from symforce_amd import _capi as capi

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        capi.hip_lib()  # loads the runtime the library uses
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), max(8, nbytes)), "hipMalloc")
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.ok(self.rt.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, nbytes):
        out = np.zeros(nbytes, np.uint8)
        self.ok(self.rt.hipMemcpy(out.ctypes.data, p, nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self, p):
        self.ok(self.rt.hipFree(p), "hipFree")


@pytest.fixture(scope="module")
def hip():
    return _Hip()


def _solve_noalloc(hip, probs, params, precision):
    t = capi.batch_pack(probs, params)
    names = ("pose_begin", "pose_soff", "match_ptr", "match_args", "odo_args", "values")
    dev = {k: hip.upload(t[k]) for k in names}
    n, stride = len(probs), params.iterations + 1
    sz_stats, sz_iter = C.sizeof(capi.OptimizationStats), C.sizeof(capi.IterationStats)
    out_vals = hip.alloc(8 * t["storage_size"])
    out_stats = hip.alloc(n * sz_stats)
    out_iters = hip.alloc(n * stride * sz_iter)
    tables = capi.BatchTables(t["n_problems"], t["total_poses"], t["storage_size"],
                              *[dev[k] for k in names])
    eps = 10 * np.finfo(np.float64).eps
    try:
        capi.batch_solve_noalloc(tables, params, precision, eps, out_vals, out_stats, out_iters)
        hip.ok(hip.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        values = hip.download(out_vals, 8 * t["storage_size"]).view(np.float64)
        stats = (capi.OptimizationStats * n).from_buffer_copy(
            hip.download(out_stats, n * sz_stats).tobytes())
        iters = (capi.IterationStats * (n * stride)).from_buffer_copy(
            hip.download(out_iters, n * stride * sz_iter).tobytes())
    finally:
        for p in list(dev.values()) + [out_vals, out_stats, out_iters]:
            hip.free(p)
    out, off = [], 0
    for i, prob in enumerate(probs):
        st = stats[i]
        sub = (capi.IterationStats * stride).from_address(C.addressof(iters) +
                                                          i * stride * sz_iter)
        out.append(dict(values=values[off:off + len(prob.values)].copy(), status=st.status,
                        failure_reason=st.failure_reason, best_index=st.best_index,
                        iterations=capi.iteration_records(sub, st.num_iterations)))
        off += len(prob.values)
    return out


def _assert_identical(got, want):
    assert len(got) == len(want)
    for a, b in zip(got, want):
        for k in ("status", "failure_reason", "best_index"):
            assert a[k] == b[k], k
        np.testing.assert_array_equal(a["values"], b["values"])
        assert a["iterations"] == b["iterations"]


@pytest.mark.parametrize("precision", [capi.PRECISION_F64, capi.PRECISION_F32_JAC])
def test_noalloc_matches_batch_solve_on_two_calls(hip, precision):
    params = capi.default_params(iterations=30)
    first = capi.gen_localization_batch(40, seed=31)
    second = capi.gen_localization_batch(7, seed=32)  # smaller: fewer poses than the first call
    assert sum(len(p.values) for p in second) < sum(len(p.values) for p in first)
    for probs in (first, second):
        want = capi.HipBatch(probs, params, precision=precision).solve()["problems"]
        got = _solve_noalloc(hip, probs, params, precision)
        _assert_identical(got, want)
        assert all(r["status"] == capi.STATUS_SUCCESS for r in got)


def test_noalloc_rejects_batches_beyond_the_scratch_size(hip):
    probs = capi.gen_localization_batch(700, seed=33, min_poses=40, max_poses=50)
    t = capi.batch_pack(probs)
    assert t["total_poses"] > capi.BATCH_NOALLOC_MAX_POSES
    tables = capi.BatchTables(t["n_problems"], t["total_poses"], t["storage_size"],
                              *([8] * 6))  # never dereferenced: the size check comes first
    with pytest.raises(RuntimeError, match="scratch"):
        capi.batch_solve_noalloc(tables, capi.default_params(), capi.PRECISION_F64, 1e-15, 8, 8)
