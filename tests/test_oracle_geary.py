"""Pins the Geary's C oracle (oracle/moran.py::geary_scatter, geary_variance): the
regular-graph identity with the pinned Moran oracle (observed and permuted), a checkerboard
closed form, and the reference test's quoted normality variances."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import moran


def _torus(side):
    """Rook adjacency on a side x side torus: symmetric, every degree 4."""
    idx = np.arange(side * side).reshape(side, side)
    rows, cols = [], []
    for dr, dc in ((0, 1), (1, 0), (0, -1), (-1, 0)):
        rows.append(idx.ravel())
        cols.append(np.roll(np.roll(idx, -dr, axis=0), -dc, axis=1).ravel())
    r, c = np.concatenate(rows), np.concatenate(cols)
    return sp.csr_matrix((np.ones(r.size), (r, c)), shape=(side * side,) * 2)


@pytest.mark.parametrize("normalise", [False, True])
def test_regular_graph_identity_with_moran(normalise):
    """On a symmetric graph with equal row sums d, sum w (x_i - x_j)^2 = 2 d denom - 2 num_I,
    so C = (n-1)/n (1 - I) for the statistic and for every permuted draw."""
    W = _torus(12)
    if normalise:
        W = moran.row_normalize(W)
    rng = np.random.default_rng(0)
    r, c = np.divmod(np.arange(144), 12)
    X = np.c_[r + c, rng.normal(size=144), np.sin(r / 3.0) * np.cos(c / 2.0)]
    perms = moran.permutations(144, 20, 3)
    C, Cs = moran.geary_scatter(X, W, perms)
    I, Is = moran.moran_scatter(X, W, perms)
    n = 144
    np.testing.assert_allclose(C, (n - 1) / n * (1 - I), rtol=0, atol=1e-12)
    np.testing.assert_allclose(Cs, (n - 1) / n * (1 - Is), rtol=0, atol=1e-12)


def test_checkerboard_closed_form():
    """Every torus edge joins opposite colours: sum w (dx)^2 = S0 and denom = n/4, so
    C = 2 (n - 1) / n."""
    W = _torus(10)
    r, c = np.divmod(np.arange(100), 10)
    C, _ = moran.geary_scatter(((r + c) % 2)[:, None], W)
    assert C[0] == pytest.approx(2 * 99 / 100, abs=1e-14)


def test_reference_quoted_variances():
    """Reference tests/test_spatial_autocorr.py:224-232: on the row-normalised Delaunay graph of
    a 22 x 22 lattice jittered by N(0, 8) (seed 1) the Geary variance is 7.052e-4 and Moran's
    6.920e-4."""
    from scipy.spatial import Delaunay
    r, c = np.divmod(np.arange(22 * 22), 22)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6]
    xy = xy + np.random.default_rng(1).normal(0, 8.0, xy.shape)
    ip, ind = Delaunay(xy).vertex_neighbor_vertices
    A = sp.csr_matrix((np.ones(ind.size), ind, ip), shape=(484, 484))
    g = moran.row_normalize(A)
    assert moran.geary_variance(g) == pytest.approx(7.052e-4, abs=5e-8)
    _, var_i = moran.analytic_pval(np.zeros(1), g)
    assert var_i == pytest.approx(6.920e-4, abs=5e-8)
