"""Shapes outside the MFMA select and the grid query, against the oracle: more than 64
neighbours, more dimensions than the select's K padding holds, spatial kNN beyond k = 32 and
outside 2-D / 3-D.  The reference serves these by argpartition (neighbors.py:66-68) and by
exact brute force (neighbors.py:350-355); here they run the exact all-pairs kernel and stay
bit-exact with the oracle, through the connectivities and the AnnData slots."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _mixture(n, d, clusters=6, seed=0):
    from metalsinglecell import synthetic
    return synthetic.gaussian_mixture_isotropic(n=n, d=d, n_clusters=clusters, seed=seed)[0]


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("n,d,k", [(3000, 50, 100), (2000, 200, 15), (1500, 200, 100),
                                   (700, 3, 65), (600, 130, 512), (1000, 126, 30)])
def test_wide_knn_bit_exact(precision, n, d, k):
    from metalsinglecell import _hip
    from oracle import knn
    X = _mixture(n, d, seed=d + k)
    idx, dist, info = _hip.knn_exact(X, k, precision=precision, return_info=True)
    ri, _, rd = knn.exact_knn_c(X, k)
    assert np.array_equal(idx.cpu().numpy(), ri)
    assert np.array_equal(dist.cpu().numpy(), rd)
    wide = k > 64 or d > (127 if precision == "fp32" else 125)
    assert (info["n_fallback"] == n and info["list_len"] == k) if wide else info["list_len"] == 64


def test_wide_knn_duplicates_keep_self_first():
    from metalsinglecell import _hip
    from oracle import knn
    X = _mixture(900, 40, seed=3)
    X = np.vstack([X, X[:100], X[:100]]).astype(np.float32)
    idx, dist = _hip.knn_exact(X, 90)
    ri, _, rd = knn.exact_knn_c(X, 90)
    idx = idx.cpu().numpy()
    assert np.array_equal(idx, ri) and np.array_equal(dist.cpu().numpy(), rd)
    assert np.array_equal(idx[:, 0], np.arange(X.shape[0]))


def test_wide_knn_limit():
    from metalsinglecell import _hip
    X = _mixture(600, 10, seed=1)
    with pytest.raises(RuntimeError, match="k=513"):
        _hip.knn_exact(X, 513)


def test_k100_distances_connectivities_and_slots():
    from metalsinglecell import _hip, pp
    from metalsinglecell._anndata import AnnData
    from oracle import fuzzy, knn
    X = _mixture(2500, 30, clusters=10, seed=5)
    X = np.vstack([X, X[:20]]).astype(np.float32)          # duplicates: zero distances kept
    n, k = X.shape[0], 100
    idx, dist = _hip.knn_exact(X, k)
    ri, _, rd = knn.exact_knn_c(X, k)
    ip, ind, dat = (t.cpu().numpy() for t in _hip.knn_distances_csr(idx, dist))
    D = knn.distances_csr(ri, rd, n)
    D.sort_indices()
    assert np.array_equal(ip, D.indptr) and np.array_equal(ind, D.indices) and np.array_equal(dat, D.data)
    (cip, cind, cdat), (sig, rho) = _hip.fuzzy_connectivities(idx, dist)
    R, rsig, rrho = fuzzy.fuzzy_simplicial_set(ri, rd)
    assert np.array_equal(sig.cpu().numpy(), rsig) and np.array_equal(rho.cpu().numpy(), rrho)
    assert np.array_equal(cip.cpu().numpy(), R.indptr)
    assert np.array_equal(cind.cpu().numpy(), R.indices)
    assert np.array_equal(cdat.cpu().numpy(), R.data)
    a = AnnData(np.zeros((n, 30), np.float32))
    a.X = X
    pp.neighbors(a, n_neighbors=k)
    assert (a.obsp["connectivities"] != R).nnz == 0
    assert (a.obsp["distances"] != D).nnz == 0
    assert a.uns["neighbors"]["params"]["n_neighbors"] == k


@pytest.mark.parametrize("k", [40, 33, 200])
def test_spatial_knn_beyond_32(k):
    from metalsinglecell import _hip
    from oracle import grid
    rng = np.random.default_rng(k)
    X = rng.uniform(0, 500, (3000, 2)).astype(np.float32)
    X[100:140] = X[:40]                                      # coincident spots
    idx, dist = _hip.spatial_knn(X, k)
    ri, rd = grid.spatial_knn_c(X, k)
    assert np.array_equal(idx.cpu().numpy(), ri) and np.array_equal(dist.cpu().numpy(), rd)


@pytest.mark.parametrize("dim", [1, 4, 5])
def test_spatial_knn_other_dimensions(dim):
    from metalsinglecell import _hip
    from oracle import grid
    X = np.random.default_rng(dim).normal(0, 100, (2000, dim)).astype(np.float32)
    X64 = X.astype(np.float64) + 0.25
    idx, dist, len64 = _hip.spatial_knn(X, 6, coords64=X64)
    ri, rd = grid.spatial_knn(X, 6)
    assert np.array_equal(idx.cpu().numpy(), ri) and np.array_equal(dist.cpu().numpy(), rd)
    rows = np.repeat(np.arange(2000), 6)
    ref = np.sqrt(((X64[rows] - X64[ri.ravel()]) ** 2).sum(1)).reshape(2000, 6)
    np.testing.assert_allclose(len64.cpu().numpy(), ref, rtol=1e-15)


def test_gr_spatial_neighbors_knn_n_neighs_40_with_libraries():
    import pandas as pd
    from metalsinglecell import gr
    from metalsinglecell._anndata import AnnData
    from oracle import grid
    rng = np.random.default_rng(11)
    codes = rng.integers(0, 2, 1600)
    xy = rng.uniform(0, 1000, (1600, 2))
    a = AnnData(np.zeros((1600, 2), np.float32))
    a.obsm["spatial"] = xy
    a.obs["lib"] = pd.Categorical(codes.astype(str))
    gr.spatial_neighbors_knn(a, n_neighs=40, library_key="lib", percentile=90.0)
    A, D = grid.section_graph_dense(xy, codes, n_neighs=40, percentile=90.0)
    assert np.array_equal(a.obsp["spatial_connectivities"].toarray(), A)
    assert np.array_equal(a.obsp["spatial_distances"].toarray(), D)
    assert sp.issparse(a.obsp["spatial_connectivities"])
