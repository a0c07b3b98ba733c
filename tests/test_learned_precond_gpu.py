"""Learned block-Jacobi correction for PCG on the reduced camera system (no reference
counterpart; the reference factorizes S directly, sparse_schur_solver.tcc:133-163).

The device network (BF16 matrix-core GEMMs, FP64 per-camera head) is checked against the CPU
reference oracle/gnn_precond.py on systems logged from the optimizer's own solves: loss and
parameter gradient; the untrained network is block Jacobi; training lowers the loss and the test
reports what the trained network gains over block Jacobi on held-out systems, with the step
error of both measured against the supernodal LDL^T solve."""
import sys
from pathlib import Path

import numpy as np
import pytest

from symforce_amd import _capi as capi

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import gnn_precond as G  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def logged():
    prob = capi.gen_ba_snavely(80, 6000, 17)
    params = capi.default_params(lambda_update_type=capi.LAMBDA_DYNAMIC, iterations=10)
    opt = capi.HipOptimizer(prob, params, capi.solver_options(
        linear_solver=capi.SOLVER_SCHUR, ordering=capi.ORDERING_NESTED_DISSECTION, use_pcg=1,
        pcg_max_iters=500, pcg_rel_tol=1e-8, pcg_preconditioner=capi.PRECOND_LEARNED))
    opt.log_systems(10)
    opt.optimize()
    assert opt.logged_systems() >= 8
    return opt


def cpu_system(opt, idx):
    row_ptr, col, blocks, rhs = opt.logged_system(idx)
    return G.System(row_ptr, col, blocks), rhs


def test_loss_and_gradient_match_cpu_reference(logged):
    opt = logged
    rng = np.random.default_rng(7)
    p = G.init_params(3)
    p["Wo"] = (0.05 * rng.standard_normal(p["Wo"].shape)).astype(np.float32)
    p["bo"] = (0.05 * rng.standard_normal(p["bo"].shape)).astype(np.float32)
    opt.preconditioner_params(G.flatten(p))
    for idx in (0, 3):
        s, _ = cpu_system(opt, idx)
        X = rng.standard_normal((s.n * 9, 4))
        loss, grad = opt.preconditioner_loss_grad(idx, X)
        x, w = G.features(s)
        loss_ref, g_ref, _ = G.loss_and_grad(p, s, x, w, X)
        gref = G.flatten(g_ref)
        assert loss == pytest.approx(loss_ref, rel=1e-4)
        assert np.linalg.norm(grad - gref) <= 1e-2 * np.linalg.norm(gref)
        for k in G.PARAM_ORDER:  # every parameter block, not just the largest
            gk, rk = G.unflatten(grad)[k], g_ref[k]
            assert np.linalg.norm(gk - rk) <= 2e-2 * np.linalg.norm(rk) + 1e-9, k


def test_untrained_network_is_block_jacobi(logged):
    opt = logged
    opt.preconditioner_params(G.flatten(G.init_params(0)))
    for idx in (1, 5):
        it_bj, err_bj = opt.eval_preconditioner(idx, 1e-8, 2000, learned=False)
        it_nn, err_nn = opt.eval_preconditioner(idx, 1e-8, 2000, learned=True)
        assert it_bj > 0 and abs(it_nn - it_bj) <= 1
        assert err_bj < 1e-5 and err_nn < 1e-5


def test_training_lowers_loss_and_reports_gain(logged):
    opt = logged
    opt.preconditioner_params(G.flatten(G.init_params(0)))
    n_sys = opt.logged_systems()
    # train on the first systems (round robin), hold out the last two
    train_opt_losses = opt.train_preconditioner(400, lr=1e-3, probes=4, seed=11,
                                                n_systems=n_sys - 2)
    assert train_opt_losses[-50:].mean() < 0.98 * train_opt_losses[:50].mean()
    gains = []
    for idx in (n_sys - 2, n_sys - 1):
        it_bj, err_bj = opt.eval_preconditioner(idx, 1e-6, 2000, learned=False)
        it_nn, err_nn = opt.eval_preconditioner(idx, 1e-6, 2000, learned=True)
        gains.append((it_bj, it_nn, err_bj, err_nn))
        assert it_nn > 0 and it_bj > 0
        assert err_nn < 1e-3 and err_bj < 1e-3
    print("PCG iterations to 1e-6 (block Jacobi, learned) and step errors vs LDL^T:", gains)
    # the trained correction must not cost iterations on held-out systems
    assert sum(g[1] for g in gains) <= 1.05 * sum(g[0] for g in gains)
