"""Writes suitesparse_patterns.json: the lower CSC patterns of A^T A + I for the six SuiteSparse
matrices the reference ships with its benchmarks
(symforce/benchmarks/matrix_multiplication/matrices/*/<name>.mtx, MatrixMarket coordinate text),
the real sparsity patterns its symbolic analysis is exercised on. Only the patterns are kept
(values are not needed for a symbolic check).  Run from the repo root (reads /root/reference):
    python tests/golden/make_suitesparse_patterns.py
"""
import json
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/symforce/benchmarks/matrix_multiplication/matrices")
NAMES = ["b1_ss", "bibd_9_3", "lp_sc105", "n3c4-b2", "rotor1", "Tina_DisCog"]
OUT = Path(__file__).resolve().parent / "suitesparse_patterns.json"


def read_mtx_pattern(path):
    lines = [ln for ln in path.read_text().splitlines() if ln.strip() and not ln.startswith("%")]
    m, n, nnz = (int(x) for x in lines[0].split()[:3])
    symmetric = "symmetric" in path.read_text().splitlines()[0]
    rows, cols = [], []
    for ln in lines[1:1 + nnz]:
        i, j = (int(x) - 1 for x in ln.split()[:2])
        rows.append(i)
        cols.append(j)
        if symmetric and i != j:
            rows.append(j)
            cols.append(i)
    A = np.zeros((m, n), bool)
    A[rows, cols] = True
    return A


def main():
    out = {}
    for name in NAMES:
        A = read_mtx_pattern(SRC / name / f"{name}.mtx").astype(np.int64)
        S = (A.T @ A) > 0
        np.fill_diagonal(S, True)
        n = S.shape[0]
        colptr, rowind = [0], []
        for c in range(n):
            r = np.nonzero(S[c:, c])[0] + c
            rowind.extend(int(x) for x in r)
            colptr.append(len(rowind))
        out[name] = {"n": n, "colptr": colptr, "rowind": rowind, "source_shape": list(A.shape)}
        print(name, A.shape, "->", n, "nnz(lower)", len(rowind))
    OUT.write_text(json.dumps(out))


if __name__ == "__main__":
    main()
