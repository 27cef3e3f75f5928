"""Pins the spatial-kNN and Moran oracles (brute force, closed forms, reference KATs)."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import grid, moran


def test_spatial_numpy_c_and_float64():
    X = np.random.default_rng(0).uniform(0, 100, (1500, 2)).astype(np.float32)
    a = grid.spatial_knn(X, 6)
    b = grid.spatial_knn_c(X, 6)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    D = np.sqrt(((X[:, None].astype(np.float64) - X[None]) ** 2).sum(-1))
    np.fill_diagonal(D, np.inf)
    ref = np.argsort(D, axis=1, kind="stable")[:, :6]
    assert np.mean([set(a[0][i]) == set(ref[i]) for i in range(1500)]) == 1.0
    np.testing.assert_allclose(a[1], np.take_along_axis(D, a[0], 1), rtol=1e-6)


def test_hex_lattice_fixed_degree_binary():
    """Reference tests/test_spatial_neighbors.py:79-85 (fixed degree, binary data)."""
    r, c = np.divmod(np.arange(400), 20)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6]
    A, D = grid.knn_graph(xy, 6)
    assert np.all(np.diff(A.indptr) == 6) and set(np.unique(A.data)) == {1.0}
    assert D.diagonal().sum() == 0


def _lattice_signals(side=25, seed=0):
    r, c = np.divmod(np.arange(side * side), side)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6]
    rng = np.random.default_rng(seed)
    X = np.c_[(r + c).astype(np.float32), ((r + c) % 2).astype(np.float32),
              rng.normal(size=side * side)].astype(np.float32)
    return xy, X


def test_moran_dense_equals_scatter_and_known_answers():
    """Reference KATs tests/test_spatial_autocorr.py:61-70."""
    xy, X = _lattice_signals()
    A, _ = grid.knn_graph(xy, 6)
    W = moran.row_normalize(A)
    perms = moran.permutations(X.shape[0], 999, 0)
    I, sims = moran.moran_dense(X, W, perms)
    I2, sims2 = moran.moran_scatter(X, W, perms[:50])
    np.testing.assert_allclose(I, I2, atol=1e-12)
    np.testing.assert_allclose(sims[:50], sims2, atol=1e-12)
    p_sim, _, _ = moran.permutation_pvals(I, sims)
    assert I[1] < 0 < I[0]
    assert p_sim[0] == pytest.approx(1 / 1000) and p_sim[1] == pytest.approx(1 / 1000)
    assert p_sim[2] > 0.05


def test_transformation_invariant_on_regular_graph():
    """Reference tests/test_spatial_autocorr.py:118-132: constant out-degree -> I unchanged."""
    xy, X = _lattice_signals(15)
    A, _ = grid.knn_graph(xy, 6)
    a, _ = moran.moran_dense(X, moran.row_normalize(A))
    b, _ = moran.moran_dense(X, sp.csr_matrix(A, dtype=np.float64))
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_graph_moments_brute_force():
    rng = np.random.default_rng(1)
    W = sp.random(30, 30, density=0.2, random_state=1, format="csr")
    Wd = W.toarray()
    s0, s1, s2 = moran.graph_moments(W)
    assert s0 == pytest.approx(Wd.sum())
    assert s1 == pytest.approx(0.5 * sum((Wd[i, j] + Wd[j, i]) ** 2 for i in range(30) for j in range(30)))
    assert s2 == pytest.approx(sum((Wd[i].sum() + Wd[:, i].sum()) ** 2 for i in range(30)))


def test_bh_definition_and_edge_cases():
    rng = np.random.default_rng(0)
    for p in [rng.uniform(size=50), rng.uniform(size=30) ** 4, np.repeat(0.03, 10),
              np.array([1.0, 1.0, 1e-12]), np.array([0.5])]:
        got = moran.bh(p)
        n = p.size
        o = np.argsort(p)
        ref = np.empty(n)
        for r in range(n):
            ref[o[r]] = min(1.0, min(p[o[j]] * n / (j + 1) for j in range(r, n)))
        np.testing.assert_allclose(got, ref, rtol=1e-12)
