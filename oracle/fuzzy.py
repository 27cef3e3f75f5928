"""UMAP fuzzy-simplicial-set connectivities oracle — TEST INFRASTRUCTURE ONLY.

The reference calls umap-learn's ``fuzzy_simplicial_set`` with precomputed kNN,
``set_op_mix_ratio=1.0`` and ``local_connectivity=1.0`` (src/metalsinglecell/neighbors.py:
742-753).  umap-learn 0.5.12 (CODE_AVAILABILITY.md:25) is not installed here, so this
restates its published algorithm (McInnes, Healy & Melville 2018; umap-learn
``smooth_knn_dist`` / ``compute_membership_strengths``) at those arguments:

1. ``rho_i`` = the first strictly positive distance of row i (rows ascend), else 0.
2. ``sigma_i`` by at most 64 bisection steps from lo=0, hi=inf, mid=1 on
   ``psum(mid) = sum_{j>=1} [d_ij - rho_i > 0 ? exp(-(d_ij - rho_i)/mid) : 1]`` toward
   ``log2(k)``, stopping when ``|psum - log2 k| < 1e-5``; float32 lo/mid/hi/psum.
3. Floor: ``sigma_i >= 1e-3 * mean(row i)`` when rho_i > 0, else ``>= 1e-3 * mean(all)``.
4. Membership ``v_ij``: 0 for j == i, 1 if d_ij - rho_i <= 0 or sigma_i == 0, else
   ``exp(-(d_ij - rho_i) / sigma_i)`` in float32.
5. ``P = (A + A^T) - A o A^T`` on the zero-eliminated COO matrix A, float32, canonical CSR.

Deliberate definitions (shared bit-for-bit with the HIP kernels): ``exp`` is the project's
float32 ``expf32`` (oracle/numerics.py) instead of umap-learn's numba ``np.exp``; the means
in step 3 are accumulated in float64.  Both move results by an ulp or so, far inside the
1e-5 tolerance, but keep GPU and oracle identical.
"""

from __future__ import annotations

import numpy as np

from .numerics import expf32

SMOOTH_K_TOLERANCE = 1e-5
MIN_K_DIST_SCALE = 1e-3


def smooth_knn_dist(dist, k, n_iter=64):
    """Per-row ``(sigma, rho)`` as float32 arrays (vectorised over rows)."""
    D = np.ascontiguousarray(dist, dtype=np.float32)
    n = D.shape[0]
    target = np.log2(float(k))
    pos = D > 0
    has = pos.any(axis=1)
    first = np.argmax(pos, axis=1)
    rho = np.where(has, D[np.arange(n), first], np.float32(0.0)).astype(np.float32)
    lo = np.zeros(n, np.float32)
    hi = np.full(n, np.inf, np.float32)
    mid = np.ones(n, np.float32)
    done = np.zeros(n, bool)
    for _ in range(n_iter):
        psum = np.zeros(n, np.float32)
        for j in range(1, D.shape[1]):
            d = (D[:, j] - rho).astype(np.float32)
            e = expf32(-(d / mid).astype(np.float32))
            psum = (psum + np.where(d > 0, e, np.float32(1.0))).astype(np.float32)
        conv = np.abs(psum.astype(np.float64) - target) < SMOOTH_K_TOLERANCE
        done |= conv
        act = ~done
        gt = act & (psum.astype(np.float64) > target)
        le = act & ~gt
        hi = np.where(gt, mid, hi)
        mid_gt = ((lo + hi) / np.float32(2.0)).astype(np.float32)
        lo = np.where(le, mid, lo)
        mid_le = np.where(hi == np.inf, mid * np.float32(2.0),
                          ((lo + hi) / np.float32(2.0))).astype(np.float32)
        mid = np.where(gt, mid_gt, np.where(le, mid_le, mid)).astype(np.float32)
        if done.all():
            break
    sigma = mid.copy()
    mean_all = float(np.sum(D, dtype=np.float64)) / D.size
    mean_row = np.zeros(n)
    for j in range(D.shape[1]):          # sequential float64 row sums, column order
        mean_row = mean_row + D[:, j].astype(np.float64)
    mean_row = mean_row / D.shape[1]
    floor = np.where(rho > 0, MIN_K_DIST_SCALE * mean_row, MIN_K_DIST_SCALE * mean_all)
    sigma = np.where(sigma.astype(np.float64) < floor, floor.astype(np.float32), sigma)
    return sigma.astype(np.float32), rho


def membership(idx, dist, sigma, rho):
    """Directed membership strengths ``v`` (n, k) float32 (step 4)."""
    D = np.asarray(dist, dtype=np.float32)
    n = D.shape[0]
    diff = (D - rho[:, None]).astype(np.float32)
    e = expf32(-(diff / sigma[:, None]).astype(np.float32))
    v = np.where((diff <= 0) | (sigma[:, None] == 0), np.float32(1.0), e).astype(np.float32)
    v[np.asarray(idx) == np.arange(n)[:, None]] = 0.0
    return v


def fuzzy_simplicial_set(idx, dist, k=None):
    """Symmetric connectivities CSR (float32) from a self-inclusive kNN."""
    import scipy.sparse as sp
    idx = np.asarray(idx)
    n, kk = idx.shape
    k = kk if k is None else k
    sigma, rho = smooth_knn_dist(dist, k)
    v = membership(idx, dist, sigma, rho)
    keep = idx.ravel() >= 0
    rows = np.repeat(np.arange(n), kk)[keep]
    A = sp.coo_matrix((v.ravel()[keep], (rows, idx.ravel()[keep])), shape=(n, n))
    A.eliminate_zeros()
    At = A.transpose()
    P = (A + At) - A.multiply(At)
    P = sp.csr_matrix(P, dtype=np.float32)
    P.eliminate_zeros()
    P.sort_indices()
    return P, sigma, rho
