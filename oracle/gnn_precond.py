"""CPU reference of the learned block-Jacobi correction for PCG on the reduced camera system.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the device network's forward pass, loss and
gradients. Nothing in the product imports it. (No reference counterpart: SymForce solves the
Schur complement directly, opt/sparse_schur_solver.tcc:133-163; this is the added learned form.)

The definition the device kernels implement, in numpy (FP64 except where stated):

* Graph: camera nodes; edges are the off-diagonal blocks of S (full symmetric block-CSR).
* Node features x_a (16): log diag(S_aa) minus its mean (9); log(1 + degree);
  log(sum_b |S_ab|_F / |S_aa|_F + 1e-12); log(max_b |S_ab|_F / |S_aa|_F + 1e-12);
  log(max diag / min diag); 1; 0; 0.
* Edge weights w_ab = |S_ab|_F / sqrt(|S_aa|_F |S_bb|_F), row-normalized (A = row-stochastic).
* Three message-passing layers h' = relu([h, A h] W + b) (widths 16 -> 64 -> 64 -> 64) and a
  linear head o = h W_o + b_o (48 outputs, 45 used). Every GEMM rounds both operands to bfloat16
  (round to nearest even) and accumulates in FP32; aggregation and bias adds are FP32.
* Per node: with t = 2 tanh(o / 20) (bounded, slope 0.1 at 0), U lower triangular with
  U_rr = exp(t_rr) and U_rc = t_rc (c < r),
  L = chol(S_aa), M_a = L^-T U U^T L^-1: symmetric positive definite for any o, and o = 0 gives
  the block-Jacobi inverse S_aa^-1.
* Loss over probe vectors X (N x P), U = M S X: L = min_c |c U - X|^2 / |X|^2
  = 1 - <U, X>^2 / (|U|^2 |X|^2). It is invariant to the scale of M (as CG is) and measures the
  spread of the eigenvalues of M S seen by the probes (1 - mean^2 / mean-square).
"""
from __future__ import annotations

import numpy as np

ND = 9
NF = 16
H = 64
NO = 48
SCALE = 0.1
TRI = [(r, c) for r in range(ND) for c in range(r + 1)]  # packing of the 45 head outputs


ROUND_BF16 = True  # False: exact FP32 operands (finite-difference checks of this module)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    if not ROUND_BF16:
        return a
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def mm(a, b):
    """bf16 x bf16 -> fp32 product (fp32 accumulation emulated in float64, rounded once)."""
    return (bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)).astype(np.float32)


class System:
    """A reduced camera system in full block-CSR storage (row-major 9x9 blocks)."""

    def __init__(self, row_ptr, col, blocks):
        self.row_ptr = np.asarray(row_ptr, np.int64)
        self.col = np.asarray(col, np.int64)
        self.blocks = np.asarray(blocks, np.float64).reshape(-1, ND, ND)
        self.n = len(self.row_ptr) - 1
        self.row = np.repeat(np.arange(self.n), np.diff(self.row_ptr))
        self.diag = np.array([self.row_ptr[a] + np.nonzero(self.col[self.row_ptr[a]:self.row_ptr[a + 1]] == a)[0][0]
                              for a in range(self.n)])

    def matvec(self, x):
        """S x for x of shape (n*9,) or (n*9, P)."""
        xs = x.reshape(self.n, ND, -1)
        y = np.zeros_like(xs)
        np.add.at(y, self.row, np.einsum("tij,tjp->tip", self.blocks, xs[self.col]))
        return y.reshape(x.shape)

    def dense(self):
        S = np.zeros((self.n * ND, self.n * ND))
        for t in range(len(self.col)):
            a, b = self.row[t], self.col[t]
            S[a * ND:(a + 1) * ND, b * ND:(b + 1) * ND] = self.blocks[t]
        return S


def features(sys: System):
    nrm = np.linalg.norm(sys.blocks.reshape(len(sys.col), -1), axis=1)
    naa = nrm[sys.diag]
    off = sys.row != sys.col
    x = np.zeros((sys.n, NF), np.float32)
    d = np.array([np.diag(sys.blocks[sys.diag[a]]) for a in range(sys.n)])
    ld = np.log(d)
    x[:, :ND] = ld - ld.mean(axis=1, keepdims=True)
    deg = np.bincount(sys.row[off], minlength=sys.n)
    osum = np.bincount(sys.row[off], weights=nrm[off], minlength=sys.n)
    omax = np.zeros(sys.n)
    np.maximum.at(omax, sys.row[off], nrm[off])
    x[:, 9] = np.log1p(deg)
    x[:, 10] = np.log(osum / naa + 1e-12)
    x[:, 11] = np.log(omax / naa + 1e-12)
    x[:, 12] = np.log(d.max(axis=1) / d.min(axis=1))
    x[:, 13] = 1.0
    w = np.where(off, nrm / np.sqrt(naa[sys.row] * naa[sys.col]), 0.0)
    rs = np.bincount(sys.row, weights=w, minlength=sys.n)
    w = np.where(rs[sys.row] > 0, w / np.where(rs[sys.row] > 0, rs[sys.row], 1.0), 0.0)
    return x, w.astype(np.float32)


def aggregate(sys, w, h, transpose=False):
    """(A h)_a = sum_b w_ab h_b, or (A^T h)_b = sum_a w_ab h_a; FP32."""
    out = np.zeros_like(h, dtype=np.float64)
    if transpose:
        np.add.at(out, sys.col, w[:, None].astype(np.float64) * h[sys.row])
    else:
        np.add.at(out, sys.row, w[:, None].astype(np.float64) * h[sys.col])
    return out.astype(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, fan_in in (("W1", 2 * NF), ("W2", 2 * H), ("W3", 2 * H)):
        p[name] = (rng.standard_normal((fan_in, H)) * np.sqrt(2.0 / fan_in)).astype(np.float32)
        p["b" + name[1]] = np.zeros(H, np.float32)
    p["Wo"] = np.zeros((H, NO), np.float32)
    p["bo"] = np.zeros(NO, np.float32)
    return p


PARAM_ORDER = ("W1", "b1", "W2", "b2", "W3", "b3", "Wo", "bo")


def flatten(p):
    return np.concatenate([p[k].ravel() for k in PARAM_ORDER]).astype(np.float32)


def unflatten(v):
    shapes = {"W1": (2 * NF, H), "b1": (H,), "W2": (2 * H, H), "b2": (H,), "W3": (2 * H, H),
              "b3": (H,), "Wo": (H, NO), "bo": (NO,)}
    out, off = {}, 0
    for k in PARAM_ORDER:
        n = int(np.prod(shapes[k]))
        out[k] = np.asarray(v[off:off + n], np.float32).reshape(shapes[k])
        off += n
    return out


def forward(p, sys, x, w):
    cache = {"h0": x}
    h = x
    for l in (1, 2, 3):
        z = np.concatenate([h, aggregate(sys, w, h)], axis=1)
        pre = mm(z, p[f"W{l}"]) + p[f"b{l}"]
        h = np.maximum(pre, 0).astype(np.float32)
        cache[f"z{l}"], cache[f"p{l}"], cache[f"h{l}"] = z, pre, h
    o = mm(h, p["Wo"]) + p["bo"]
    cache["o"] = o
    return o, cache


def precond_blocks(sys, o):
    """M_a = L^-T U U^T L^-1 per node (n, 9, 9) and the factors used by the backward pass."""
    n = sys.n
    M = np.zeros((n, ND, ND))
    Linv = np.zeros((n, ND, ND))
    U = np.zeros((n, ND, ND))
    for a in range(n):
        L = np.linalg.cholesky(sys.blocks[sys.diag[a]])
        Li = np.linalg.inv(L)
        Ua = np.zeros((ND, ND))
        for k, (r, c) in enumerate(TRI):
            t = 2.0 * np.tanh(float(o[a, k]) / 20.0)
            Ua[r, c] = np.exp(t) if r == c else t
        K = Ua.T @ Li
        M[a] = K.T @ K
        Linv[a], U[a] = Li, Ua
    return M, Linv, U


def loss_and_grad(p, sys, x, w, X):
    """Loss and parameter gradients for probe matrix X (n*9, P)."""
    o, c = forward(p, sys, x, w)
    M, Linv, U = precond_blocks(sys, o)
    Y = sys.matvec(X).reshape(sys.n, ND, -1)
    Xs = X.reshape(sys.n, ND, -1)
    Uy = np.einsum("aij,ajp->aip", M, Y)
    Z = float(np.sum(X * X))
    A = float(np.sum(Uy * Xs))
    B = float(np.sum(Uy * Uy))
    loss = 1.0 - A * A / (B * Z)
    E = (A / B) * Uy - Xs
    G = 2.0 * A / (B * Z) * np.einsum("aip,ajp->aij", E, Y)
    dU = np.einsum("aij,ajk,alk,alm->aim", Linv, G + np.transpose(G, (0, 2, 1)), Linv, U)
    do = np.zeros_like(o, dtype=np.float32)
    for k, (r, cc) in enumerate(TRI):
        dt = 0.1 * (1.0 - np.tanh(o[:, k].astype(np.float64) / 20.0) ** 2)
        do[:, k] = dt * dU[:, r, cc] * (U[:, r, r] if r == cc else 1.0)
    g = {}
    g["Wo"] = mm(c["h3"].T, do)
    g["bo"] = do.sum(axis=0).astype(np.float32)
    dh = mm(do, p["Wo"].T)
    for l in (3, 2, 1):
        dpre = np.where(c[f"p{l}"] > 0, dh, 0).astype(np.float32)
        g[f"W{l}"] = mm(c[f"z{l}"].T, dpre)
        g[f"b{l}"] = dpre.sum(axis=0).astype(np.float32)
        if l > 1:
            dz = mm(dpre, p[f"W{l}"].T)
            dh = (dz[:, :H] + aggregate(sys, w, dz[:, H:], transpose=True)).astype(np.float32)
    return loss, g, M


def pcg_iterations(sys, M, b, tol, max_iters=10000):
    """Iterations of PCG with block preconditioner M (n, 9, 9) to ||r|| <= tol ||b||."""
    x = np.zeros_like(b)
    r = b.copy()
    z = np.einsum("aij,aj->ai", M, r.reshape(sys.n, ND)).ravel()
    p_ = z.copy()
    rz = r @ z
    b2 = b @ b
    for k in range(max_iters):
        if r @ r <= tol * tol * b2:
            return k
        q = sys.matvec(p_)
        alpha = rz / (p_ @ q)
        x += alpha * p_
        r -= alpha * q
        z = np.einsum("aij,aj->ai", M, r.reshape(sys.n, ND)).ravel()
        rz_new = r @ z
        p_ = z + rz_new / rz * p_
        rz = rz_new
    return max_iters
