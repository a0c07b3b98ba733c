"""include_jacobians and check_derivatives on the device, against the oracle's sparse
linearizer (which builds the Jacobian from the reference's triplets, linearizer.cc:475-538):

* the Jacobian's pattern (stats.jacobian_sparsity) bit-equal to the oracle's, on every problem
  family the registry covers (poses, rotations, BA with Huber and the linear camera with Cauchy,
  the localization factors, the GNC Barron factor);
* the best linearization's Jacobian values equal the oracle's at the best Values (1e-9), and
  under debug_stats every record's jacobian_values equal the oracle's linearization at the
  record's Values;
* check_derivatives (the device compares every factor body's Jacobian with central differences
  through the retraction, derivative_checker.h:32-73) passes on all of them and is refused
  without include_jacobians, as the reference's Optimizer asserts (optimizer.tcc:39)."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from problem_builders import gnc_problem, gnc_reference_samples

pytestmark = pytest.mark.gpu
# The rotation log's f(m) = acos(m) / sqrt(1 - m^2) at small relative rotations carries ~1e-9
# relative rounding noise in any FP64 evaluation of the reference's formula (the oracle's
# included), which its derivative amplifies: entries agree to 1e-8 of the largest.
TOL = 1e-8

def retyped(prob, old, new, edit):
    """The problem with every factor of type `old` rewritten as type `new` (its arguments through
    edit(args, add_key)); add_key(letter, value) appends a Scalar key and returns its index."""
    letters, subs, supers = list(prob.letters), list(prob.subs), list(prob.supers)
    types, values = list(prob.types), list(prob.values)

    def add_key(letter, value):
        letters.append(ord(letter))
        subs.append(np.iinfo(np.int64).min)
        supers.append(np.iinfo(np.int64).min)
        types.append(capi.TYPE_SCALAR)
        values.append(value)
        return len(letters) - 1

    ftypes, fptr, fargs = [], [0], []
    for f, t in enumerate(prob.ftypes):
        args = list(prob.fargs[prob.fptr[f]:prob.fptr[f + 1]])
        if t == old:
            t, args = new, edit(args, add_key)
        ftypes.append(t)
        fargs.extend(args)
        fptr.append(len(fargs))
    return capi.Problem(letters, subs, supers, types, values, ftypes, fptr, fargs,
                        prob.opt_keys, prob.info)


def snavely_plain():
    # SNAVELY_HUBER (pose, intrinsics, point, pixel, k, eps) -> SNAVELY (pose, ..., pixel, eps)
    return retyped(capi.gen_ba_snavely(6, 60, 3), capi.FACTOR_SNAVELY_HUBER, capi.FACTOR_SNAVELY,
                   lambda a, add: a[:4] + a[5:])


def between_cauchy():
    # BETWEEN_POSE3 (a, b, a_T_b, sqrt_info, eps) -> BETWEEN_POSE3_CAUCHY (..., c, eps)
    cache = {}

    def edit(a, add):
        if "c" not in cache:
            cache["c"] = add("c", 0.5)
        return a[:4] + [cache["c"]] + a[4:]
    return retyped(capi.gen_pose_graph(200, 30, 5), capi.FACTOR_BETWEEN_POSE3,
                   capi.FACTOR_BETWEEN_POSE3_CAUCHY, edit)


FAMILIES = {
    "pose_smoothing": lambda: capi.gen_pose_smoothing(10, 42),
    "rot_smoothing": lambda: capi.gen_rot_smoothing(10, 42),
    "rot_frozen": lambda: capi.gen_rot_frozen(3, 42),
    "ba_circle_cauchy": lambda: capi.gen_ba_circle(6, 80, 3),
    "ba_snavely_huber": lambda: capi.gen_ba_snavely(6, 60, 3),
    "ba_snavely_plain": snavely_plain,
    "pose_graph": lambda: capi.gen_pose_graph(200, 30, 5),
    "pose_graph_cauchy": between_cauchy,
    "localization": lambda: capi.gen_localization(8, 10, 0.5, 3),
    "gnc_barron": lambda: gnc_problem(gnc_reference_samples())[0],
}


@pytest.mark.parametrize("family", sorted(FAMILIES))
def test_jacobian_pattern_and_values_match_oracle(family):
    prob = FAMILIES[family]()
    params = capi.default_params(iterations=8, include_jacobians=1, debug_stats=1,
                                 check_derivatives=1)
    opt = capi.HipOptimizer(prob, params)
    res = opt.optimize()
    op = ow.OracleProblem(prob, include_jacobians=True)
    cp, ri = opt.jacobian_sparsity(op.N)
    ocp, ori = op.jacobian_structure()
    np.testing.assert_array_equal(cp, ocp)
    np.testing.assert_array_equal(ri, ori)
    nnz = ri.size
    best = op.linearize(res["values"])["jacobian_values"]
    got = opt.best_jacobian(nnz)
    np.testing.assert_allclose(got, best, rtol=0, atol=TOL * max(1.0, np.abs(best).max()))
    for i in range(len(res["iterations"])):
        vals = np.zeros(prob.values.size)
        capi.check(capi.hip_lib().symhip_optimizer_debug_record(
            opt.h, i, None, capi.ptr(vals, capi.C.c_double), None))
        want = op.linearize(vals)["jacobian_values"]
        got = opt.debug_jacobian(i, nnz)
        np.testing.assert_allclose(got, want, rtol=0, atol=TOL * max(1.0, np.abs(want).max()))


def test_check_derivatives_requires_include_jacobians():
    with pytest.raises(RuntimeError, match="include_jacobians"):
        capi.HipOptimizer(capi.gen_pose_smoothing(10, 42),
                          capi.default_params(check_derivatives=1))


def test_no_jacobian_without_include_jacobians():
    opt = capi.HipOptimizer(capi.gen_pose_smoothing(10, 42), capi.default_params())
    opt.optimize()
    with pytest.raises(RuntimeError, match="include_jacobians"):
        opt.jacobian_sparsity(60)


def test_jacobian_on_the_schur_path():
    prob = capi.gen_ba_snavely(6, 60, 3)
    params = capi.default_params(iterations=6, include_jacobians=1, check_derivatives=1)
    opts = capi.solver_options(linear_solver=capi.SOLVER_SCHUR)
    opt = capi.HipOptimizer(prob, params, opts)
    res = opt.optimize()
    op = ow.OracleProblem(prob, include_jacobians=True)
    cp, ri = opt.jacobian_sparsity(op.N)
    np.testing.assert_array_equal(ri, op.jacobian_structure()[1])
    best = op.linearize(res["values"])["jacobian_values"]
    np.testing.assert_allclose(opt.best_jacobian(ri.size), best, rtol=0,
                               atol=TOL * np.abs(best).max())
