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


def test_sigma_floor_uses_the_row_mean_when_rho_is_positive():
    """Four neighbours tie at rho, so psum >= 4 > log2(8) for every sigma: the bisection halves
    mid 64 times and the floor 1e-3 * mean(row i) (column 0 included) must bind.  A second row
    with far larger distances makes the global mean differ from the row mean, so taking the
    wrong mean fails; a third row that converges normally must keep its bisection sigma."""
    k = 8
    idx = np.array([[i] + [(i + j) % 3 + 3 * j for j in range(1, k)] for i in range(3)])
    dist = np.array([[0, 2, 2, 2, 2, 5, 6, 7],
                     [0, 100, 150, 200, 250, 300, 350, 400],
                     [0, 1, 1.5, 2, 2.5, 3, 3.5, 4]], np.float32)
    sigma, rho = fuzzy.smooth_knn_dist(dist, k)
    assert rho[0] == 2.0 and rho[1] == 100.0 and rho[2] == 1.0
    assert sigma[0] == np.float32(1e-3 * (0 + 2 * 4 + 5 + 6 + 7) / 8)
    mean_all = float(dist.astype(np.float64).sum()) / dist.size
    assert sigma[0] != np.float32(1e-3 * mean_all)
    # rows 1 and 2 converge above their floors: sum exp(-(d - rho)/sigma) = log2 k
    for i in (1, 2):
        assert sigma[i] > 1e-3 * dist[i].mean()
        s = np.exp(-np.maximum(dist[i, 1:] - rho[i], 0) / np.float64(sigma[i])).sum()
        assert abs(s - 3.0) < 1e-4
    v = fuzzy.membership(idx, dist, sigma, rho)
    assert np.all(v[0, 1:5] == 1.0)              # the tied nearest neighbours weigh 1
    assert np.all(v[0, 5:] < 1e-30)              # exp(-(3..5)/3.25e-3) underflows


def test_sigma_floor_uses_the_global_mean_when_rho_is_zero():
    """A row of exact duplicates has no positive distance: rho = 0, every term is 1, psum =
    k-1 > log2 k, and the floor is 1e-3 * the mean of ALL n*k distances."""
    k = 6
    dist = np.array([[0, 0, 0, 0, 0, 0],
                     [0, 1, 2, 3, 4, 5],
                     [0, 10, 20, 30, 40, 50]], np.float32)
    sigma, rho = fuzzy.smooth_knn_dist(dist, k)
    assert rho[0] == 0.0
    mean_all = float(dist.astype(np.float64).sum()) / dist.size
    assert sigma[0] == np.float32(1e-3 * mean_all)
    assert sigma[0] != np.float32(0.0)


def test_rho_skips_zero_distance_duplicates():
    """Duplicates sit at distance 0 after the point itself; rho is the first STRICTLY positive
    distance (column 3 here), not column 1, and the duplicates get weight 1."""
    k = 9
    idx = np.arange(9)[None, :]
    dist = np.array([[0, 0, 0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0]], np.float32)
    sigma, rho = fuzzy.smooth_knn_dist(dist, k)
    assert rho[0] == np.float32(1.5)
    v = fuzzy.membership(idx, dist, sigma, rho)
    assert v[0, 0] == 0.0 and np.all(v[0, 1:4] == 1.0)
    # the five remaining terms must make up log2(9) - 3 at the converged sigma
    s = np.exp(-(dist[0, 4:].astype(np.float64) - 1.5) / np.float64(sigma[0])).sum()
    assert abs(3.0 + s - np.log2(9)) < 1e-4
