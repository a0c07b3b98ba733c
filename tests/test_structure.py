"""Bit-exact checks of the host symbolic layer (state order, hessian_lower pattern, per-factor
storage maps) against the oracle's restatement of BuildInitialLinearization. CPU only."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
import problem_builders as pb

CASES = {
    "pose_smoothing": lambda: capi.gen_pose_smoothing(10, 42),
    "rot_smoothing": lambda: capi.gen_rot_smoothing(10, 42),
    "rot_frozen": lambda: capi.gen_rot_frozen(3, 42),
    "robot3d": lambda: pb.robot3d_problem()[0],
    "ba_circle": lambda: capi.gen_ba_circle(6, 60, 3),
    "ba_snavely": lambda: capi.gen_ba_snavely(12, 80, 5),
    "pose_graph": lambda: capi.gen_pose_graph(2000, 150, 7),
    "localization": lambda: capi.gen_localization(12, 15, 0.3, 9),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_structure_matches_oracle(name):
    prob = CASES[name]()
    hp = capi.HipProblem(prob)
    op = ow.OracleProblem(prob)
    assert (hp.N, hp.M, hp.nnz) == (op.N, op.M, op.nnz)
    cp, ri = hp.hessian_structure()
    ocp, ori = op.hessian_structure()
    np.testing.assert_array_equal(cp, ocp)
    np.testing.assert_array_equal(ri, ori)
    np.testing.assert_array_equal(hp.col_starts(), op.hessian_col_starts())
    k, o = hp.state_keys()
    ok, oo = op.state_keys()
    np.testing.assert_array_equal(k, ok)
    np.testing.assert_array_equal(o, oo)


def test_pose_smoothing_sizes_match_reference():
    """test/symforce_optimizer_test.cc:477-506: H 60x60 with 534 nonzeros, residual 66."""
    hp = capi.HipProblem(capi.gen_pose_smoothing(10, 42))
    assert (hp.N, hp.M, hp.nnz) == (60, 66, 534)


def test_unoptimized_key_rejected():
    prob = capi.gen_pose_smoothing(4, 1)
    # add an optimized key that no factor touches
    b = pb.ProblemBuilder()
    b.letters, b.subs, b.supers = list(prob.letters), list(prob.subs), list(prob.supers)
    b.types, b.values = list(prob.types), list(prob.values)
    b.ftypes, b.fptr, b.fargs = list(prob.ftypes), list(prob.fptr), list(prob.fargs)
    extra = b.key("z", type_=capi.TYPE_VECTOR3, value=[0, 0, 0])
    b.opt_keys = [0, 1, 2, 3, extra]
    with pytest.raises(RuntimeError, match="not optimized by any factor"):
        capi.HipProblem(b.build())
