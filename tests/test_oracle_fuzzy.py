"""Pins the connectivities oracle by closed forms, invariants and a float64 bisection."""

import numpy as np

from oracle import fuzzy, knn


def _knn(n=800, d=10, k=15, seed=0):
    X = np.random.default_rng(seed).standard_normal((n, d)).astype(np.float32)
    idx, _, dist = knn.exact_knn_c(X, k)
    return idx, dist


def test_invariants():
    idx, dist = _knn()
    P, sigma, rho = fuzzy.fuzzy_simplicial_set(idx, dist)
    assert (abs(P - P.T) > 0).nnz == 0
    assert P.data.min() > 0 and P.data.max() <= 1.0
    assert P.diagonal().sum() == 0
    v = fuzzy.membership(idx, dist, sigma, rho)
    # nearest positive-distance neighbour has directed weight 1
    assert np.all(v[:, 1] == 1.0)
    # sum_{j>=1} exp(-(d - rho)/sigma) = log2 k at convergence
    s = np.exp(-np.maximum(dist[:, 1:] - rho[:, None], 0) / sigma[:, None].astype(np.float64)).sum(1)
    assert np.max(np.abs(s - np.log2(15))) < 1e-4


def test_sigma_matches_float64_bisection():
    """Independent float64 bisection with the true exponential: sigma within 1e-5 relative."""
    idx, dist = _knn(300, 8, 12, seed=3)
    sigma, rho = fuzzy.smooth_knn_dist(dist, 12)
    target = np.log2(12)
    for i in range(0, 300, 7):
        d = dist[i, 1:].astype(np.float64) - rho[i]
        lo, hi = 0.0, 1e6
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            ps = np.where(d > 0, np.exp(-d / mid), 1.0).sum()
            lo, hi = (lo, mid) if ps > target else (mid, hi)
        assert abs(sigma[i] - mid) <= 1e-5 * mid + 1e-7


def test_special_cases():
    # k = 2: the single neighbour is rho itself -> psum = 1 = log2(2) at mid = 1 -> weight 1
    idx = np.array([[0, 1], [1, 0], [2, 1]])
    dist = np.array([[0, 2.0], [0, 2.0], [0, 5.0]], np.float32)
    P, sigma, rho = fuzzy.fuzzy_simplicial_set(idx, dist)
    assert np.allclose(rho, [2, 2, 5]) and np.allclose(sigma, 1.0)
    assert P[0, 1] == 1.0 and P[1, 0] == 1.0 and P[2, 1] == 1.0 and P[1, 2] == 1.0
    # all distances equal: every (d - rho) = 0 -> weight 1
    n, k = 30, 6
    idx = np.array([[i] + [(i + j) % n for j in range(1, k)] for i in range(n)])
    dist = np.c_[np.zeros(n), np.full((n, k - 1), 3.0)].astype(np.float32)
    v = fuzzy.membership(idx, dist, *fuzzy.smooth_knn_dist(dist, k))
    assert np.all(v[:, 1:] == 1.0) and np.all(v[:, 0] == 0.0)


def test_symmetrisation_closed_form():
    """P_ij = a + b - ab with a = A_ij, b = A_ji, on a hand-built 3-node case."""
    idx = np.array([[0, 1, 2], [1, 2, 0], [2, 0, 1]])
    dist = np.array([[0, 1.0, 2.0], [0, 1.5, 3.0], [0, 0.5, 4.0]], np.float32)
    P, sigma, rho = fuzzy.fuzzy_simplicial_set(idx, dist)
    v = fuzzy.membership(idx, dist, sigma, rho)
    A = np.zeros((3, 3), np.float32)
    for i in range(3):
        for j in range(3):
            A[i, idx[i, j]] = v[i, j]
    ref = (A + A.T) - A * A.T
    assert np.array_equal(P.toarray(), ref)


def test_duplicates_no_nan():
    """Reference tests/test_postfix_fixes.py:197-207."""
    rng = np.random.default_rng(0)
    X = rng.standard_normal((60, 10)).astype(np.float32)
    X = np.vstack([X, X[:5]]).astype(np.float32)
    idx, _, dist = knn.exact_knn(X, 10)
    P, _, _ = fuzzy.fuzzy_simplicial_set(idx, dist)
    assert not np.isnan(P.data).any()
