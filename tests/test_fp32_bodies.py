"""The FP32 factor bodies of the precision forms (no reference counterpart: the reference evaluates
its generated functions in the optimizer's Scalar).

The FP32 bodies form the quantities that cancel -- relative rotations q_a^-1 q_b, the point in a
camera frame, measured-minus-predicted terms -- in FP64 from the FP64 Values and round them once
(relative Lie-group deltas), so their Jacobians keep FP32's relative accuracy far from the
origin. The bar: on problems whose coordinates are large against their relative motion (the
1000-camera random-walk bundle adjustment, the 100 000-node pose graph), the FP32 body is more
accurate than any FP32 evaluation that starts from the absolute Values rounded to FP32 (measured
here by the FP64 body on rounded inputs, a lower bound on such an evaluation's error), and its
error stays within a few FP32 roundings of the FP64 Jacobian. Host evaluation of the same
bodies the device runs; no GPU needed."""
import numpy as np
import pytest

from symforce_amd import _capi as capi


def factor_args(prob, f):
    out = []
    for k in prob.fargs[prob.fptr[f]:prob.fptr[f + 1]]:
        o = prob.storage_offsets[k]
        out.append(prob.values[o:o + capi.STORAGE_DIM[int(prob.types[k])]])
    return out


def errors(prob, n, seed=0):
    rng = np.random.default_rng(seed)
    by_type = {}
    for f in rng.choice(prob.n_factors, size=min(n, prob.n_factors), replace=False):
        ft = int(prob.ftypes[f])
        a = factor_args(prob, f)
        r64, J64 = capi.factor_linearize_host(ft, a, capi.PRECISION_F64)
        r32, J32 = capi.factor_linearize_host(ft, a, capi.PRECISION_F32_JAC)
        rounded = [x.astype(np.float32).astype(np.float64) for x in a]
        _, Jrd = capi.factor_linearize_host(ft, rounded, capi.PRECISION_F64)
        nrm = np.linalg.norm(J64)
        if nrm == 0:
            continue
        e = by_type.setdefault(ft, ([], []))
        e[0].append(np.linalg.norm(J32 - J64) / nrm)
        e[1].append(np.linalg.norm(Jrd - J64) / nrm)
    return {ft: (np.array(a), np.array(b)) for ft, (a, b) in by_type.items()}


@pytest.mark.parametrize("name,gen,types", [
    ("ba_snavely_1000", lambda: capi.gen_ba_snavely(1000, 20000, 42),
     [capi.FACTOR_SNAVELY_HUBER]),
    ("pose_graph_100k", lambda: capi.gen_pose_graph(100000, 30000, 42),
     [capi.FACTOR_BETWEEN_POSE3, capi.FACTOR_BETWEEN_POSE3_CAUCHY]),
])
def test_fp32_bodies_beat_rounded_absolute_inputs(name, gen, types):
    prob = gen()
    errs = errors(prob, 1500)
    for ft in types:
        rel, rounded = errs[ft]
        print(f"\n{name} type {ft}: FP32 body median {np.median(rel):.2e} max {rel.max():.2e}; "
              f"rounded inputs median {np.median(rounded):.2e} max {rounded.max():.2e}")
        assert np.median(rel) < np.median(rounded) / 5
        assert rel.max() < rounded.max() / 5
        assert np.median(rel) < 5e-7 and rel.max() < 1e-5


def test_fp32_bodies_localization_and_residuals():
    """The localization factors (the batched FP32 form's bodies) to a few FP32 roundings, and the
    FP32 residuals of all bodies to FP32 accuracy of the measurement-scale quantities."""
    for prob in (capi.gen_localization(50, 40, 0.3, 3), capi.gen_ba_circle(20, 300, 4)):
        errs = errors(prob, 2000)
        for ft, (rel, _) in errs.items():
            assert rel.max() < 2e-5, (ft, rel.max())
    prob = capi.gen_pose_graph(20000, 3000, 5)
    for f in range(0, prob.n_factors, 97):
        ft = int(prob.ftypes[f])
        a = factor_args(prob, f)
        r64, _ = capi.factor_linearize_host(ft, a, capi.PRECISION_F64)
        r32, _ = capi.factor_linearize_host(ft, a, capi.PRECISION_F32_JAC)
        assert np.abs(r32 - r64).max() <= 1e-5 * max(1.0, np.abs(r64).max())


def test_fp64_host_bodies_match_the_oracle():
    """The FP64 bodies (the expressions the FP32 bodies' deltas come from) against the oracle's
    per-factor linearization (forward-mode dual numbers through the restated residuals)."""
    import oracle_wrap as ow
    for prob in (capi.gen_ba_snavely(30, 400, 2), capi.gen_pose_graph(3000, 400, 11),
                 capi.gen_localization(30, 25, 0.3, 5), capi.gen_ba_circle(8, 150, 3)):
        for f in range(0, prob.n_factors, max(1, prob.n_factors // 300)):
            ft = int(prob.ftypes[f])
            a = factor_args(prob, f)
            r, J = capi.factor_linearize_host(ft, a, capi.PRECISION_F64)
            (res, jac, _, _), _ = ow.linearize_factor(ft, a)
            D, T = J.shape
            want_J = jac[:D * T].reshape(T, D).T  # oracle: column-major D x T
            # (residuals of noiseless priors are rounding noise of the coordinates)
            np.testing.assert_allclose(r, res[:D], rtol=1e-10, atol=1e-10)
            np.testing.assert_allclose(J, want_J, rtol=1e-9, atol=1e-9 * max(1, np.abs(J).max()))
