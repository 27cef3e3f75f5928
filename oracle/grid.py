"""Exact spatial kNN and fixed-radius oracle (2-D / 3-D) — TEST INFRASTRUCTURE ONLY.

Restates what ``gr.spatial_neighbors_knn`` builds (src/metalsinglecell/gr.py:174-225 on
``neighbors._knn_grid``, neighbors.py:333-417): a directed graph with ``n_neighs`` edges per
observation, self excluded, neighbours ordered by ``(float32 distance, index)``
(gr.py:190-193).  Connectivities are 1.0 in float32 with columns in that order
(gr.py:218-220); distances are float64 edge lengths from float64 coordinates (gr.py:32-34,
214-217).

The float32 distance is ``sqrt(fl(fl(dx*dx) + fl(dy*dy)))`` (plus ``fl(dz*dz)`` in 3-D),
``dx = fl(x_j - x_i)``, every operation separately rounded.  Direct differences do not
cancel at large coordinates — the failure the reference pins in
tests/test_spatial_neighbors.py:377-386 (two spots 273 apart at offsets up to 2e5).
"""

from __future__ import annotations

import numpy as np


def spatial_knn(coords, k, block=512):
    """Brute force: ``(idx int64 (n, k), dist float32 (n, k))`` with self excluded."""
    X = np.ascontiguousarray(coords, dtype=np.float32)
    n, d = X.shape
    if not 1 <= k < n:
        raise ValueError("need 1 <= k < n")
    out_i = np.empty((n, k), np.int64)
    out_d = np.empty((n, k), np.float32)
    cols = np.arange(n)
    for s in range(0, n, block):
        Q = X[s:s + block]
        acc = None
        for t in range(d):
            diff = (X[None, :, t] - Q[:, t, None]).astype(np.float32)
            sq = (diff * diff).astype(np.float32)
            acc = sq if acc is None else (acc + sq).astype(np.float32)
        dist = np.sqrt(acc).astype(np.float32)
        for r in range(Q.shape[0]):
            i = s + r
            drow = dist[r].copy()
            drow[i] = np.inf
            kth = np.partition(drow, k - 1)[k - 1]
            cand = np.flatnonzero(drow <= kth)
            cand = cand[cand != i]
            o = np.lexsort((cols[cand], drow[cand]))[:k]
            out_i[i] = cand[o]
            out_d[i] = drow[cand[o]]
    return out_i, out_d


def spatial_knn_c(coords, k, threads=0):
    """Same definition in C (knn_ref.c ``oracle_spatial_knn``)."""
    import ctypes
    from .knn import _lib
    X = np.ascontiguousarray(coords, dtype=np.float32)
    n, d = X.shape
    lib = _lib()
    lib.oracle_spatial_knn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int]
    out_i = np.empty((n, k), np.int64)
    out_d = np.empty((n, k), np.float32)
    rc = lib.oracle_spatial_knn(X.ctypes.data, n, d, k, out_i.ctypes.data,
                                out_d.ctypes.data, int(threads))
    if rc != 0:
        raise RuntimeError(f"oracle_spatial_knn failed ({rc})")
    return out_i, out_d


def knn_graph(coords, n_neighs=6, exact=None):
    """``(connectivities, distances)`` CSR pair as written by ``spatial_neighbors_knn``."""
    import scipy.sparse as sp
    X = np.asarray(coords)
    n = X.shape[0]
    idx, _ = spatial_knn(X, n_neighs) if n <= 20_000 else spatial_knn_c(X, n_neighs)
    ex = np.asarray(X if exact is None else exact, dtype=np.float64)
    rows = np.repeat(np.arange(n), n_neighs)
    cols = idx.ravel()
    diff = ex[rows] - ex[cols]
    dist = np.sqrt((diff * diff).sum(1))
    indptr = np.arange(n + 1, dtype=np.int64) * n_neighs
    A = sp.csr_matrix((np.ones(dist.size, np.float32), cols, indptr), shape=(n, n))
    D = sp.csr_matrix((dist, cols, indptr), shape=(n, n))
    return A, D


def radius_neighbors(coords, r_max, r_min=0.0, block=512):
    """All pairs with ``r_min^2 <= d2 <= r_max^2``, self excluded (neighbors.py:255-330 and its
    kernel :199-252): ``(indptr int64, indices int64, dist float32)``, columns ascending.

    ``d2`` is the separately rounded float32 sum of this module; the bounds are float64
    squares rounded to float32, as the reference passes them to its kernel (:306-307)."""
    X = np.ascontiguousarray(coords, dtype=np.float32)
    n, d = X.shape
    r2hi = np.float32(float(r_max) ** 2)
    r2lo = np.float32(float(r_min) ** 2)
    counts = np.zeros(n, np.int64)
    cols, dists = [], []
    for s in range(0, n, block):
        Q = X[s:s + block]
        acc = None
        for t in range(d):
            diff = (X[None, :, t] - Q[:, t, None]).astype(np.float32)
            sq = (diff * diff).astype(np.float32)
            acc = sq if acc is None else (acc + sq).astype(np.float32)
        keep = (acc >= r2lo) & (acc <= r2hi)
        keep[np.arange(Q.shape[0]), np.arange(s, s + Q.shape[0])] = False
        r, c = np.nonzero(keep)
        counts[s:s + Q.shape[0]] = np.bincount(r, minlength=Q.shape[0])
        cols.append(c)
        dists.append(np.sqrt(acc[r, c]).astype(np.float32))
    indptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=indptr[1:])
    return indptr, np.concatenate(cols).astype(np.int64), np.concatenate(dists)


def radius_graph(coords, r_max, r_min=0.0, exact=None):
    """``(connectivities, distances)`` CSR pair as written by ``spatial_neighbors_radius``
    (gr.py:228-252) before the percentile / diagonal / transform tail."""
    import scipy.sparse as sp
    X = np.asarray(coords)
    n = X.shape[0]
    indptr, cols, _ = radius_neighbors(X, r_max, r_min)
    ex = np.asarray(X if exact is None else exact, dtype=np.float64)
    rows = np.repeat(np.arange(n), np.diff(indptr))
    diff = ex[rows] - ex[cols]
    dist = np.sqrt((diff * diff).sum(1))
    A = sp.csr_matrix((np.ones(dist.size, np.float32), cols, indptr), shape=(n, n))
    D = sp.csr_matrix((dist, cols, indptr.copy()), shape=(n, n))
    return A, D


def section_graph_dense(xy64, codes=None, n_neighs=None, radius=None, percentile=None,
                        set_diag=False, transform=None):
    """Dense ``(connectivities, distances)`` of a spatial graph after squidpy's tail, built
    section by section as the reference does (gr.py:81-147, 198-252): per section (cells with
    one library code, in code order) the directed kNN (``n_neighs``) or radius graph of that
    section alone; edges longer than the section's ``percentile`` of edge lengths removed
    from both matrices; ``set_diag`` puts 1.0 on the connectivity diagonal; the distances
    keep no diagonal and no zero; ``transform`` "spectral" scales by the column degrees
    (``A / sqrt(deg_i deg_j)``, 0 where a degree is 0), "cosine" is the cosine similarity of
    the rows.  Dense float64 n x n matrices: a different formulation from the product's
    sparse assembly, for a few thousand cells."""
    X64 = np.asarray(xy64, dtype=np.float64)
    n = X64.shape[0]
    codes = np.zeros(n, np.int64) if codes is None else np.asarray(codes)
    A = np.zeros((n, n))
    D = np.zeros((n, n))
    for c in np.unique(codes):
        ids = np.flatnonzero(codes == c)
        sub = X64[ids]
        m = ids.size
        E = np.zeros((m, m), bool)
        if n_neighs is not None:
            k = min(int(n_neighs), m - 1)
            if k > 0:
                idx, _ = spatial_knn(sub.astype(np.float32), k)
                E[np.repeat(np.arange(m), k), idx.ravel()] = True
        else:
            lo, hi = (0.0, float(radius)) if np.isscalar(radius) else (min(radius), max(radius))
            ip, cols, _ = radius_neighbors(sub.astype(np.float32), hi, lo)
            E[np.repeat(np.arange(m), np.diff(ip)), cols] = True
        L = np.sqrt(((sub[:, None, :] - sub[None, :, :]) ** 2).sum(-1))
        if percentile is not None and E.any():
            E &= L <= np.percentile(L[E], percentile)
        As = E.astype(np.float64)
        if set_diag:
            np.fill_diagonal(As, 1.0)
        Ds = np.where(E, L, 0.0)
        np.fill_diagonal(Ds, 0.0)
        A[np.ix_(ids, ids)] = As
        D[np.ix_(ids, ids)] = Ds
    if transform == "spectral":
        deg = A.sum(axis=0)
        inv = np.where(deg > 0, 1.0 / np.sqrt(np.where(deg > 0, deg, 1.0)), 0.0)
        A = A * inv[:, None] * inv[None, :]
    elif transform == "cosine":
        norm = np.sqrt((A * A).sum(axis=1))
        U = np.where(norm[:, None] > 0, A / np.where(norm > 0, norm, 1.0)[:, None], 0.0)
        A = U @ U.T
    return A, D
