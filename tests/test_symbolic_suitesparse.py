"""The symbolic analysis (AnalyzeSparsityPattern, sparse_cholesky_solver.tcc:26-106) on real
sparsity patterns: A^T A + I of the six SuiteSparse matrices the reference ships with its
benchmarks (tests/golden/suitesparse_patterns.json, written by make_suitesparse_patterns.py).

The elimination tree and the column counts of L for the permutation the solver chooses are held
to a brute-force symbolic elimination of the dense boolean pattern of P A P^T (fill: eliminating
column k connects every pair of rows below it that column k reaches), independent of the oracle's
restated etree algorithm, and to the oracle as well; natural, minimum-degree and
nested-dissection orderings. No GPU needed (the symbolic layer runs on the host)."""
import json
from pathlib import Path

import numpy as np
import pytest

from symforce_amd import _capi as capi
import oracle_wrap as ow

PATTERNS = json.loads((Path(__file__).resolve().parent / "golden" /
                       "suitesparse_patterns.json").read_text())


def brute_force(n, colptr, rowind, inv_perm):
    """Parent and strictly-lower column counts of L for P A P^T by dense boolean elimination."""
    S = np.zeros((n, n), bool)
    for c in range(n):
        r = rowind[colptr[c]:colptr[c + 1]]
        S[r, c] = True
        S[c, r] = True
    B = S[np.ix_(inv_perm, inv_perm)]
    L = np.tril(B, -1)
    parent = np.full(n, -1, np.int64)
    for k in range(n):
        rows = np.nonzero(L[k + 1:, k])[0] + k + 1
        if rows.size:
            parent[k] = rows[0]
            L[np.ix_(rows, rows)] |= np.tril(np.ones((rows.size, rows.size), bool), -1)
    return parent, L.sum(axis=0)


@pytest.mark.parametrize("name", sorted(PATTERNS))
@pytest.mark.parametrize("ordering", [capi.ORDERING_NATURAL, capi.ORDERING_MIN_DEGREE,
                                      capi.ORDERING_NESTED_DISSECTION])
def test_symbolic_on_suitesparse_patterns(name, ordering):
    m = PATTERNS[name]
    n = m["n"]
    cp = np.array(m["colptr"], np.int32)
    ri = np.array(m["rowind"], np.int32)
    ch = capi.HipCholesky(n, cp, ri, ordering=ordering)
    inv_perm = ch.inv_permutation()
    assert sorted(inv_perm.tolist()) == list(range(n))
    parent, counts = ch.etree()
    bf_parent, bf_counts = brute_force(n, cp, ri, inv_perm)
    np.testing.assert_array_equal(counts, bf_counts)
    np.testing.assert_array_equal(np.where(parent < 0, -1, parent), bf_parent)
    oc = ow.OracleCholesky(n, cp, ri, None if ordering == capi.ORDERING_NATURAL else inv_perm)
    np.testing.assert_array_equal(np.diff(oc.factor()["colptr"]), bf_counts)
    # the supernodal panels hold every structural nonzero of L
    assert ch.info()["panel"] >= int(bf_counts.sum()) + n


def test_minimum_degree_reduces_fill():
    """Minimum degree never fills more than the natural order on these patterns by more than a
    few percent (it halves rotor1's fill). Nested dissection is printed beside it: on these small
    patterns it is no fill reducer (lp_sc105: 1355 against 874 natural), its purpose is the
    elimination tree's depth on large graphs."""
    for name, m in PATTERNS.items():
        n, cp, ri = m["n"], np.array(m["colptr"], np.int32), np.array(m["rowind"], np.int32)
        fill = {o: capi.HipCholesky(n, cp, ri, ordering=o).etree()[1].sum()
                for o in (capi.ORDERING_NATURAL, capi.ORDERING_MIN_DEGREE,
                          capi.ORDERING_NESTED_DISSECTION)}
        print(name, fill)
        assert fill[capi.ORDERING_MIN_DEGREE] <= 1.05 * fill[capi.ORDERING_NATURAL] + 5
