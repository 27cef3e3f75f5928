"""Pins the bbknn oracle (oracle/bbknn.py): one batch reduces to the pinned exact kNN, the
per-batch lists match a float64 k-d tree, the reference's tiny-batch case, and the trimming
rule against a hand-worked graph and its closed-form characterisation."""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.spatial import cKDTree

from oracle import bbknn, knn


def _data(n=300, d=12, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, d)).astype(np.float32)


def test_single_batch_is_the_exact_knn():
    X = _data()
    idx, dist, K = bbknn.bbknn_knn(X, np.zeros(X.shape[0], int), 7)
    ri, _, rd = knn.exact_knn(X, 7)
    assert K == 7 and np.array_equal(idx, ri) and np.array_equal(dist, rd)


def test_per_batch_lists_match_kdtree():
    """Each cell's kk_b neighbours from batch b are the k-d tree's (float64) nearest of batch b;
    the combined row is ascending and holds exactly kk_b cells of each batch."""
    X = _data(400, 8, 1)
    rng = np.random.default_rng(1)
    batch = rng.choice(np.array(["x", "a", "m"]), 400, p=[0.5, 0.3, 0.2])
    idx, dist, K = bbknn.bbknn_knn(X, batch, 4)
    codes, nb = bbknn.batch_codes(batch)
    assert K == 12 and nb == 3
    assert (np.diff(dist, axis=1) >= 0).all()
    X64 = X.astype(np.float64)
    for b in range(nb):
        rows = np.flatnonzero(codes == b)
        _, nn = cKDTree(X64[rows]).query(X64, 4)
        got = np.sort(np.where(codes[idx] == b, idx, -1), axis=1)[:, -4:]
        assert np.array_equal(got, np.sort(rows[nn], axis=1))
    # own batch: the cell itself is in its list at distance 0
    assert ((idx == np.arange(400)[:, None]).sum(axis=1) == 1).all()


def test_tiny_batch_reference_case():
    """Reference tests/test_postfix_fixes.py:186-194: a 2-cell batch with k=3 takes 2 columns."""
    rng = np.random.default_rng(0)
    X = rng.standard_normal((60, 10)).astype(np.float32)
    batch = np.array(["a"] * 40 + ["b"] * 2 + ["c"] * 18)
    D, conn, K = bbknn.bbknn(X, batch, neighbors_within_batch=3)
    assert K == 8 and conn.shape == (60, 60) and conn.nnz > 0
    assert (conn != conn.T).nnz == 0
    assert D.nnz == 60 * 7 and (D.diagonal() == 0).all()


def test_trim_hand_example():
    """Row 0 holds 4 entries, trim 2: its threshold is its 2nd largest (0.8), so (0,1)=0.5
    and (0,3)=0.3 go; every other row is short (threshold 0) but loses the same entries
    through row 0's threshold on the column pass."""
    A = np.zeros((5, 5), np.float32)
    for (i, j, v) in [(0, 1, 0.5), (0, 2, 0.9), (0, 3, 0.3), (0, 4, 0.8), (1, 2, 0.4), (3, 4, 0.7)]:
        A[i, j] = A[j, i] = v
    T = bbknn.trim_graph(sp.csr_matrix(A), 2).toarray()
    E = A.copy()
    E[0, 1] = E[1, 0] = E[0, 3] = E[3, 0] = 0.0
    assert np.array_equal(T, E)


@pytest.mark.parametrize("trim", [1, 3, 6, 50])
def test_trim_closed_form(trim):
    """The literal row / transpose / row procedure keeps (i, j) iff v != 0 and v >= both
    rows' thresholds (the trim-th largest of rows longer than trim, else 0)."""
    rng = np.random.default_rng(trim)
    n = 80
    A = sp.random(n, n, density=0.15, random_state=rng, format="csr", dtype=np.float32)
    A = sp.csr_matrix(A.maximum(A.T))
    A.data = np.round(A.data, 2).astype(np.float32)   # ties on purpose
    A.eliminate_zeros()
    thr = np.zeros(n, np.float32)
    for i in range(n):
        r = A.data[A.indptr[i]:A.indptr[i + 1]]
        if r.size > trim:
            thr[i] = np.sort(r)[::-1][trim - 1]
    C = A.tocoo()
    keep = (C.data != 0) & (C.data >= thr[C.row]) & (C.data >= thr[C.col])
    E = sp.csr_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=(n, n))
    T = bbknn.trim_graph(A, trim)
    assert (T != E).nnz == 0 and T.nnz == E.nnz
