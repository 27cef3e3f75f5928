"""HIP fixed-radius neighbours (msc_spatial_radius_count / _fill) against the CPU oracle:
identical CSR structure, float32 distances and float64 lengths."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _check(X, r_max, r_min=0.0, exact=None):
    from metalsinglecell import _hip
    from oracle import grid
    out = _hip.spatial_radius(X, r_max, r_min, coords64=exact)
    ip, ind, dist = (t.cpu().numpy() for t in out[:3])
    rip, rind, rdist = grid.radius_neighbors(X, r_max, r_min)
    assert np.array_equal(ip, rip)
    assert np.array_equal(ind, rind)
    assert np.array_equal(dist, rdist)
    if exact is not None:
        _, D = grid.radius_graph(X, r_max, r_min, exact=exact)
        assert np.array_equal(out[3].cpu().numpy(), D.data)
    return ip, ind


@pytest.mark.parametrize("case", range(3))
def test_radius_sklearn_datasets(case):
    """The reference's sklearn comparison datasets (tests/test_spatial_neighbors.py:363-375)."""
    rng = np.random.RandomState(0)
    data = [(rng.uniform(0, 100, (3000, 2)).astype(np.float32), 6.0),
            ((rng.normal(size=(3000, 2)) * [30, 3]).astype(np.float32), 4.0),
            (rng.uniform(0, 50, (2000, 3)).astype(np.float32), 7.0)]
    X, r = data[case]
    _check(X, r, exact=X.astype(np.float64))


def test_radius_lattice_boundaries_and_intervals():
    """An exact hex lattice with the radius on the first shell (ties at the cutoff) and
    intervals that cut between shells."""
    r, c = np.divmod(np.arange(900), 30)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6].astype(np.float32)
    for lo, hi in [(0.0, 100.0), (0.0, 99.99), (100.0, 100.0), (130.0, 200.0), (0.0, 0.0)]:
        _check(xy, hi, lo)


def test_radius_duplicates_large_offsets_and_3d():
    rng = np.random.default_rng(3)
    X = rng.uniform(0, 1000, (4000, 2)).astype(np.float32)
    X[2000:2100] = X[:100]                      # coincident pairs
    _check(X, 25.0)
    _check(X, 0.0)                              # only the coincident pairs
    base = np.array([[9237.0, 15447.0], [9510.0, 15447.0], [9237.0, 15720.0]])
    for off in (0.0, 2e4, 2e5):
        Y = np.ascontiguousarray(np.r_[base, rng.uniform(0, 400, (500, 2)) + base[0]] + off, dtype=np.float32)
        _check(Y, 273.0, exact=Y.astype(np.float64))
    Z = np.concatenate([rng.normal(m, 5.0, (800, 3)) for m in (0.0, 40.0, 90.0)]).astype(np.float32)
    _check(Z, 6.5, 1.0)


def test_radius_elongated_and_degenerate_boxes():
    """A long thin strip forces cell coarsening; a flat 3-D set and a single repeated point
    exercise zero extents."""
    rng = np.random.default_rng(4)
    strip = np.c_[rng.uniform(0, 1e6, 5000), rng.uniform(0, 1.0, 5000)].astype(np.float32)
    _check(strip, 300.0)
    flat = np.c_[rng.uniform(0, 100, (2000, 2)), np.zeros(2000)].astype(np.float32)
    _check(flat, 4.0)
    same = np.ones((300, 2), np.float32)
    ip, ind = _check(same, 1.0)
    assert (np.diff(ip) == 299).all()


def test_gr_spatial_neighbors_radius_slots():
    from metalsinglecell import gr
    from metalsinglecell._anndata import AnnData
    from oracle import grid
    rng = np.random.default_rng(5)
    r, c = np.divmod(np.arange(400), 20)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6] + rng.normal(0, 2.0, (400, 2))
    a = AnnData(np.zeros((400, 2), np.float32))
    a.obsm["spatial"] = xy
    gr.spatial_neighbors_radius(a, radius=(130.0, 200.0))
    A, D = grid.radius_graph(xy.astype(np.float32), 200.0, 130.0, exact=xy)
    assert (a.obsp["spatial_connectivities"] != A).nnz == 0
    assert (a.obsp["spatial_distances"] != D).nnz == 0
    assert a.uns["spatial_neighbors"]["params"]["radius"] == (130.0, 200.0)
    res = gr.spatial_neighbors_radius(a, radius=200.0, transform="spectral", set_diag=True, copy=True)
    assert (abs(res.connectivities - res.connectivities.T) > 1e-12).nnz == 0
