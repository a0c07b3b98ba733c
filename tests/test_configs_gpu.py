"""The benchmark's other configurations as parity cases at their real sizes.

* 20 cameras on a circle, 1000 landmarks, Cauchy loss, prior on the first pose, LM with the
  sparse Cholesky solver (the repo-CPU-path BA configuration);
* pose-graph SLAM: 100 000 Pose3 nodes, odometry + loop closures (1 % outliers, Cauchy), the
  supernodal LDL^T with nested-dissection ordering.

Both are held to the oracle's LM trajectory computed with the same elimination order (the
reference orders with METIS; the ordering itself is not a parity surface, see test_cholesky)."""
import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow
from test_optimizer_gpu import assert_same_trajectory

pytestmark = pytest.mark.gpu


def nd_permutation(prob):
    """The elimination order the optimizer's nested dissection picks (nodes = optimized keys)."""
    hp = capi.HipProblem(prob)
    cp, ri = hp.hessian_structure()
    n = len(cp) - 1
    _, offsets = hp.state_keys()
    node_ptr = np.append(offsets, n).astype(np.int32)
    return capi.HipCholesky(n, cp, ri, node_ptr=node_ptr,
                            ordering=capi.ORDERING_NESTED_DISSECTION).inv_permutation()


def test_circle_ba_20_cameras_1000_landmarks():
    prob = capi.gen_ba_circle(20, 1000, 42)
    params = capi.default_params(iterations=50)
    opts = capi.solver_options(linear_solver=capi.SOLVER_CHOLESKY,
                               ordering=capi.ORDERING_NESTED_DISSECTION)
    hip = capi.HipOptimizer(prob, params, opts).optimize()
    ora = ow.OracleProblem(prob).optimize(params, ordering=nd_permutation(prob))
    assert_same_trajectory(hip, ora, rtol=1e-8)


def test_pose_graph_100k_nodes():
    prob = capi.gen_pose_graph(100000, 30000, 42)
    params = capi.default_params(iterations=12)
    opts = capi.solver_options(linear_solver=capi.SOLVER_CHOLESKY,
                               ordering=capi.ORDERING_NESTED_DISSECTION)
    hip = capi.HipOptimizer(prob, params, opts).optimize()
    ora = ow.OracleProblem(prob).optimize(params, ordering=nd_permutation(prob))
    assert_same_trajectory(hip, ora, rtol=1e-8)
    first, best = hip["iterations"][0]["new_error"], hip["iterations"][hip["best_index"]]["new_error"]
    assert best < 0.9 * first
