"""Batch-balanced kNN oracle — TEST INFRASTRUCTURE ONLY.

Restates the reference's ``neighbors.bbknn`` (src/metalsinglecell/neighbors.py:491-559) and
``_trim_graph`` (:465-488):

1. Batches are the present categories in ``np.unique`` order (:510-511); batch b's cells
   are its rows in ascending order (:515).
2. For every cell and every batch b, the ``kk_b = min(k, n_b)`` nearest cells of batch b
   (:520, :530-537), concatenated in batch order (:542-543).
3. The concatenation is sorted by distance (:544-546) and handed to the fuzzy simplicial
   set with ``n_neighbors = K = sum_b kk_b`` (:548-553).
4. ``distances``: the K entries per row as a CSR with the diagonal and zeros removed
   (:554-557).
5. ``_trim_graph(conn, 10 K or trim)``: row i's threshold is its trim-th largest value when
   the row is longer than trim (else 0), applied to rows, then to the transpose's rows.

Deliberate definitions shared with the HIP kernels (documented deviations):
* distances are the exact kNN distances of oracle/knn.py (float32 sum of squared differences
  in coordinate order, each operation rounded), not the reference's Gram expansion
  ``|x|^2 + |y|^2 - 2 x.y`` (:532-533); within a batch the order key is ``(j != i, d2, j)``;
* the combined sort is stable (ties keep the batch order) where the reference's
  ``np.argsort`` leaves the order of equal distances unspecified (:544).
"""

from __future__ import annotations

import numpy as np

from . import fuzzy


def batch_codes(batch):
    cats, codes = np.unique(np.asarray(batch), return_inverse=True)
    return codes.ravel().astype(np.int64), int(cats.size)


def _sqdist(Q, C):
    """float32 squared distances (rows of Q x rows of C), coordinate order, each op rounded."""
    acc = np.zeros((Q.shape[0], C.shape[0]), dtype=np.float32)
    for t in range(Q.shape[1]):
        diff = C[None, :, t] - Q[:, t, None]
        acc += diff * diff
    return acc


def per_batch_knn(X, codes, n_batches, k, block=512):
    """Per-batch lists, concatenated in batch order: ``(idx int64 (n, K), d2 float32 (n, K))``."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n = X.shape[0]
    parts_i, parts_d = [], []
    for b in range(n_batches):
        rows = np.flatnonzero(codes == b)
        kk = min(k, rows.size)
        oi = np.empty((n, kk), np.int64)
        od = np.empty((n, kk), np.float32)
        for s in range(0, n, block):
            q = np.arange(s, min(n, s + block))
            D = _sqdist(X[q], X[rows])
            for r, i in enumerate(q):
                not_self = (rows != i).astype(np.int8)
                o = np.lexsort((rows, D[r], not_self))[:kk]
                oi[i] = rows[o]
                od[i] = D[r, o]
        parts_i.append(oi)
        parts_d.append(od)
    return np.concatenate(parts_i, axis=1), np.concatenate(parts_d, axis=1)


def bbknn_knn(X, batch, k):
    """Combined lists sorted by float32 distance (stable): ``(idx int64, dist float32, K)``."""
    codes, nb = batch_codes(batch)
    idx, d2 = per_batch_knn(X, codes, nb, k)
    dist = np.sqrt(d2).astype(np.float32)
    order = np.argsort(dist, axis=1, kind="stable")
    return np.take_along_axis(idx, order, axis=1), np.take_along_axis(dist, order, axis=1), idx.shape[1]


def distances_csr(idx, dist, n):
    """neighbors.py:554-557: all K entries, diagonal set to 0, zeros eliminated."""
    import scipy.sparse as sp
    K = idx.shape[1]
    rows = np.repeat(np.arange(n), K)
    D = sp.csr_matrix((dist.ravel(), (rows, idx.ravel())), shape=(n, n))
    D = D.tolil()
    D.setdiag(0)
    D = D.tocsr()
    D.eliminate_zeros()
    D.sort_indices()
    return D


def trim_graph(conn, trim):
    """``_trim_graph`` (neighbors.py:465-488), literally: thresholds from the rows, then
    rows, transpose, rows again with the same thresholds."""
    import scipy.sparse as sp
    cnts = sp.csr_matrix(conn, copy=True)
    if trim is None or trim <= 0:
        return cnts
    vals = np.zeros(cnts.shape[0])
    for i in range(cnts.shape[0]):
        row = cnts.data[cnts.indptr[i]:cnts.indptr[i + 1]]
        if row.shape[0] > trim:
            vals[i] = row[np.argsort(row, kind="stable")[-trim]]
    for _ in range(2):
        for i in range(cnts.shape[0]):
            row = cnts.data[cnts.indptr[i]:cnts.indptr[i + 1]]
            row[row < vals[i]] = 0
        cnts.eliminate_zeros()
        cnts = cnts.T.tocsr()
    cnts.sort_indices()
    return cnts


def bbknn(X, batch, neighbors_within_batch=3, trim=None):
    """``(distances, connectivities, K)`` scipy CSR, as ``neighbors.bbknn`` (n_pcs applied by
    the caller)."""
    n = X.shape[0]
    idx, dist, K = bbknn_knn(X, batch, neighbors_within_batch)
    conn, _, _ = fuzzy.fuzzy_simplicial_set(idx, dist)
    conn = trim_graph(conn, 10 * K if trim is None else trim)
    return distances_csr(idx, dist, n), conn, K
