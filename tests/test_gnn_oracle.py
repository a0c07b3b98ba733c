"""The CPU reference of the learned preconditioner (oracle/gnn_precond.py) pinned on its own:
finite differences of its loss against its analytic gradient (with exact FP32 operands), the
untrained network reproducing the block-Jacobi inverse, and the exact inverse giving zero loss."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import gnn_precond as G  # noqa: E402


def ring_system(n=12, seed=0):
    rng = np.random.default_rng(seed)
    nb = {a: sorted({(a + d) % n for d in (-2, -1, 0, 1, 2)}) for a in range(n)}
    B = {}
    for a in range(n):
        for b in nb[a]:
            if a != b and (b, a) not in B:
                B[(a, b)] = 0.3 * rng.standard_normal((9, 9))
                B[(b, a)] = B[(a, b)].T
        R = rng.standard_normal((9, 9))
        B[(a, a)] = R @ R.T + 9 * np.eye(9)
    row_ptr, col = [0], []
    for a in range(n):
        col += nb[a]
        row_ptr.append(len(col))
    blocks = np.array([B[(a, b)] for a in range(n) for b in nb[a]])
    return G.System(row_ptr, col, blocks)


@pytest.fixture
def exact_fp32():
    G.ROUND_BF16 = False
    yield
    G.ROUND_BF16 = True


def test_gradient_matches_finite_differences(exact_fp32):
    s = ring_system()
    x, w = G.features(s)
    rng = np.random.default_rng(1)
    p = G.init_params(2)
    p["Wo"] = (0.05 * rng.standard_normal(p["Wo"].shape)).astype(np.float32)
    X = rng.standard_normal((s.n * 9, 3))
    _, g, _ = G.loss_and_grad(p, s, x, w, X)
    checked = 0
    for name in ("Wo", "bo", "W3", "b3", "W2", "W1", "b1"):
        for _ in range(6):
            idx = tuple(int(rng.integers(0, d)) for d in p[name].shape)
            h = 1e-3
            lp = G.loss_and_grad({**p, name: _bump(p[name], idx, h)}, s, x, w, X)[0]
            lm = G.loss_and_grad({**p, name: _bump(p[name], idx, -h)}, s, x, w, X)[0]
            fd = (lp - lm) / (2 * h)
            assert fd == pytest.approx(float(g[name][idx]), rel=2e-2, abs=1e-7), (name, idx)
            checked += abs(fd) > 1e-7
    assert checked >= 10


def _bump(a, idx, h):
    b = a.copy()
    b[idx] += h
    return b


def test_untrained_network_is_block_jacobi():
    s = ring_system(seed=3)
    x, w = G.features(s)
    o, _ = G.forward(G.init_params(0), s, x, w)
    assert np.all(o == 0)
    M, _, _ = G.precond_blocks(s, o)
    for a in range(s.n):
        np.testing.assert_allclose(M[a], np.linalg.inv(s.blocks[s.diag[a]]), rtol=1e-10, atol=1e-14)


def test_loss_is_scale_invariant():
    s = ring_system(seed=5)
    x, w = G.features(s)
    X = np.random.default_rng(2).standard_normal((s.n * 9, 3))
    p = G.init_params(0)
    l0 = G.loss_and_grad(p, s, x, w, X)[0]
    s2 = G.System(s.row_ptr, s.col, s.blocks * 7.0)  # M scales by 1/7, S by 7: same M S
    x2, w2 = G.features(s2)
    assert G.loss_and_grad(p, s2, x2, w2, X)[0] == pytest.approx(l0, rel=1e-12)
    assert 0.0 <= l0 < 1.0


def test_loss_vanishes_for_the_exact_inverse_on_a_block_diagonal_system():
    s = ring_system(seed=4)
    off = s.row != s.col
    s.blocks[off] = 0  # block-diagonal S: block Jacobi is the exact inverse
    x, w = G.features(s)
    X = np.random.default_rng(0).standard_normal((s.n * 9, 2))
    loss, _, _ = G.loss_and_grad(G.init_params(0), s, x, w, X)
    assert loss < 1e-20
