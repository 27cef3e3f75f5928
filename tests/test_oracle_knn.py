"""Pins the exact-kNN oracle: two implementations, fp64 brute force, sklearn, reference KATs."""

import numpy as np
import pytest

from oracle import knn


def test_numpy_and_c_oracles_agree():
    X = np.random.default_rng(0).standard_normal((500, 20)).astype(np.float32)
    for k in (1, 15, 33, 50):
        a = knn.exact_knn(X, k)
        b = knn.exact_knn_c(X, k)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_matches_float64_brute_force_and_sklearn():
    """Reference recipe tests/test_postfix_fixes.py:108-121 (500 x 20, seed 0, k 15/33/50):
    the exact oracle must reach recall 1 against float64 brute force and sklearn."""
    from sklearn.neighbors import NearestNeighbors
    X = np.random.default_rng(0).standard_normal((500, 20)).astype(np.float32)
    D = np.sqrt(((X[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1))
    for k in (15, 33, 50):
        idx, _, dist = knn.exact_knn(X, k)
        ref = np.argsort(D, axis=1, kind="stable")[:, :k]
        assert np.mean([len(set(idx[i]) & set(ref[i])) / k for i in range(500)]) == 1.0
        np.testing.assert_allclose(dist, np.take_along_axis(D, idx, 1), rtol=1e-5, atol=1e-6)
        nn = NearestNeighbors(n_neighbors=k).fit(X.astype(np.float64))
        _, si = nn.kneighbors(X.astype(np.float64))
        assert np.mean([len(set(idx[i]) & set(si[i])) / k for i in range(500)]) == 1.0


def test_order_rule_and_self_first_with_duplicates():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((40, 5)).astype(np.float32)
    X = np.vstack([X, X[:4], X[:4]]).astype(np.float32)
    idx, d2, dist = knn.exact_knn(X, 6)
    assert np.array_equal(idx[:, 0], np.arange(X.shape[0]))
    assert np.all(d2[:, 0] == 0)
    # rows ascending by (d2, index) after column 0
    for r in range(X.shape[0]):
        key = list(zip(d2[r, 1:], idx[r, 1:]))
        assert key == sorted(key)
    # a duplicate's twin comes right after self at distance 0, lower index first
    assert idx[0, 1] == 40 and dist[0, 1] == 0 and idx[40, 1] == 0


def test_large_coordinate_known_answer_273():
    """Reference tests/test_spatial_neighbors.py:377-386."""
    base = np.array([[9237.0, 15447.0], [9510.0, 15447.0], [9237.0, 15720.0]])
    for off in (0.0, 20_000.0, 200_000.0):
        X = np.ascontiguousarray(base + off, dtype=np.float32)
        _, _, d = knn.exact_knn(X, 2)
        assert abs(float(d[0, 1]) - 273.0) < 1e-3


def test_distances_csr_layout():
    X = np.random.default_rng(2).standard_normal((50, 4)).astype(np.float32)
    X[10] = X[3]
    idx, _, dist = knn.exact_knn(X, 5)
    D = knn.distances_csr(idx, dist, 50)
    assert D.has_sorted_indices and np.all(np.diff(D.indptr) == 4)
    assert D[3, 10] == 0 and 10 in D.indices[D.indptr[3]:D.indptr[4]]   # explicit zero kept
