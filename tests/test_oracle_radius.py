"""Pins the fixed-radius oracle (oracle/grid.py::radius_neighbors): the reference's own
sklearn comparison (tests/test_spatial_neighbors.py:363-375, same three datasets), symmetry
and the cutoff on the reference's jittered lattice (:89-103), and interval edge cases."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import grid


@pytest.mark.parametrize("case", range(3))
def test_radius_matches_sklearn_exactly(case):
    from sklearn.neighbors import NearestNeighbors
    rng = np.random.RandomState(0)
    data = [(rng.uniform(0, 100, (3000, 2)).astype(np.float32), 6.0),
            ((rng.normal(size=(3000, 2)) * [30, 3]).astype(np.float32), 4.0),
            (rng.uniform(0, 50, (2000, 3)).astype(np.float32), 7.0)]
    X, r = data[case]
    indptr, cols, dist = grid.radius_neighbors(X, r)
    ours = sp.csr_matrix((np.ones_like(dist), cols, indptr), shape=(len(X),) * 2)
    ref = NearestNeighbors(radius=r).fit(X).radius_neighbors_graph(mode="connectivity")
    ref.setdiag(0)
    ref.eliminate_zeros()
    assert (abs(ours - ref) > 0).nnz == 0
    assert (dist <= np.float32(r) * (1 + 1e-6)).all()


def _lattice(side, jitter, seed=0):
    r, c = np.divmod(np.arange(side * side), side)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6]
    return xy + np.random.default_rng(seed).normal(0, jitter, xy.shape)


def test_radius_symmetric_and_interval():
    """Reference lattice test: radius 200 is symmetric with lengths <= 200; the interval
    (130, 200) drops the first shell and keeps lengths inside it."""
    xy = _lattice(20, 2.0)
    A, D = grid.radius_graph(xy, 200.0, exact=xy)
    assert (abs(A - A.T) > 0).nnz == 0 and D.data.max() <= 200.0 + 1e-6
    Ab, Db = grid.radius_graph(xy, 200.0, 130.0, exact=xy)
    assert 0 < Db.nnz < D.nnz
    assert Db.data.min() >= 130.0 - 1e-6 and Db.data.max() <= 200.0 + 1e-6


def test_radius_zero_keeps_only_coincident_points():
    X = np.array([[0, 0], [1, 1], [0, 0], [2, 2], [1, 1], [0, 0]], np.float32)
    indptr, cols, dist = grid.radius_neighbors(X, 0.0)
    assert indptr.tolist() == [0, 2, 3, 5, 5, 6, 8]
    assert cols.tolist() == [2, 5, 4, 0, 5, 1, 0, 2] and (dist == 0).all()
