"""HIP grid spatial kNN against the brute-force oracle (bit-exact)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(X, k, **kw):
    from metalsinglecell import _hip
    out = _hip.spatial_knn(X, k, return_info=True, **kw)
    return [o.cpu().numpy() if hasattr(o, "cpu") else o for o in out]


def _lattice(side, jitter=0.0, seed=0):
    r, c = np.divmod(np.arange(side * side), side)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6]
    if jitter:
        xy = xy + np.random.default_rng(seed).normal(0, jitter, xy.shape)
    return xy


@pytest.mark.parametrize("k", [1, 6, 11, 20, 32])
def test_spatial_uniform(k):
    from oracle import grid
    X = np.random.default_rng(0).uniform(0, 100, (5000, 2)).astype(np.float32)
    idx, dist, info = _run(X, k)
    ri, rd = grid.spatial_knn_c(X, k)
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd), info


@pytest.mark.parametrize("jitter", [0.0, 1.0])
def test_spatial_hex_lattice_ties(jitter):
    """Exact lattice: ties everywhere; order must be (distance, index)."""
    from oracle import grid
    X = _lattice(40, jitter).astype(np.float32)
    idx, dist, _ = _run(X, 6)
    ri, rd = grid.spatial_knn(X, 6)
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)


def test_spatial_clusters_and_3d():
    from oracle import grid
    rng = np.random.default_rng(1)
    X = np.vstack([rng.normal(0, 1, (3000, 2)), rng.normal(1000, 50, (3000, 2)),
                   [[5e4, 5e4]]]).astype(np.float32)        # sparse outlier: many shells
    idx, dist, info = _run(X, 6)
    ri, rd = grid.spatial_knn_c(X, 6)
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)
    X3 = rng.uniform(0, 50, (4000, 3)).astype(np.float32)
    idx, dist, _ = _run(X3, 10)
    ri, rd = grid.spatial_knn_c(X3, 10)
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)


def test_spatial_large_coordinates_273():
    """Reference KAT tests/test_spatial_neighbors.py:377-386."""
    base = np.array([[9237.0, 15447.0], [9510.0, 15447.0], [9237.0, 15720.0]])
    for off in (0.0, 20_000.0, 200_000.0):
        X = np.ascontiguousarray(base + off)
        idx, dist, len64, _ = _run(X.astype(np.float32), 1, coords64=X)
        assert abs(float(dist[0, 0]) - 273.0) < 1e-3
        assert abs(float(len64[0, 0]) - 273.0) < 1e-9


def test_spatial_duplicates():
    from oracle import grid
    rng = np.random.default_rng(2)
    X = rng.uniform(0, 10, (500, 2)).astype(np.float32)
    X = np.vstack([X, X[:50], X[:50]]).astype(np.float32)
    idx, dist, _ = _run(X, 6)
    ri, rd = grid.spatial_knn(X, 6)
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)
