"""HIP batch-balanced kNN (msc_bbknn), its distances CSR, connectivities and trimming against
the CPU oracle: identical indices, distances and graphs."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _mixture(n, d, seed):
    from metalsinglecell import synthetic
    X, lab = synthetic.gaussian_mixture_isotropic(n=n, d=d, n_clusters=6, seed=seed)
    return X, lab


def _same_csr(A, B):
    A, B = sp.csr_matrix(A), sp.csr_matrix(B)
    A.sort_indices()
    B.sort_indices()
    return (np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
            and np.array_equal(A.data, B.data))


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("nb,k", [(1, 15), (3, 3), (5, 6)])
def test_bbknn_knn_matches_oracle(precision, nb, k):
    from metalsinglecell import _hip
    from oracle import bbknn
    X, _ = _mixture(2400, 30, nb)
    rng = np.random.default_rng(nb)
    batch = rng.integers(0, nb, X.shape[0])
    idx, dist, info = _hip.bbknn(X, batch, nb, k, precision=precision, return_info=True)
    ri, rd, K = bbknn.bbknn_knn(X, batch, k)
    assert idx.shape[1] == K
    assert np.array_equal(idx.cpu().numpy(), ri)
    assert np.array_equal(dist.cpu().numpy(), rd)
    print(f"nb={nb} k={k} {precision}: fallback lists {info['n_fallback']}")


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_bbknn_brute_force_lists(precision):
    """Huge offsets with tiny spreads defeat both rankings' certificates in the select passes:
    those (batch, cell) lists are brute-forced, and the result still equals the oracle's."""
    from metalsinglecell import _hip
    from oracle import bbknn
    rng = np.random.default_rng(17)
    centers = rng.normal(0, 1e6, (10, 8))
    X = (centers[rng.integers(0, 10, 1800)] + rng.normal(0, 1e-2, (1800, 8))).astype(np.float32)
    batch = rng.integers(0, 3, X.shape[0])
    idx, dist, info = _hip.bbknn(X, batch, 3, 5, precision=precision, return_info=True)
    ri, rd, _ = bbknn.bbknn_knn(X, batch, 5)
    assert info["n_fallback"] > 0
    assert np.array_equal(idx.cpu().numpy(), ri)
    assert np.array_equal(dist.cpu().numpy(), rd)


def test_bbknn_uneven_and_tiny_batches():
    """Batches of 1, 2, 150 and 2000 cells (padding inside query blocks, kk < k), batch codes
    interleaved in the input order."""
    from metalsinglecell import _hip
    from oracle import bbknn
    X, _ = _mixture(2153, 20, 7)
    rng = np.random.default_rng(7)
    batch = np.array(["big"] * 2000 + ["mid"] * 150 + ["two"] * 2 + ["one"])
    rng.shuffle(batch)
    codes, nb = bbknn.batch_codes(batch)
    for precision in ("fp32", "bf16"):
        idx, dist = _hip.bbknn(X, codes, nb, 4, precision=precision)
        ri, rd, K = bbknn.bbknn_knn(X, batch, 4)
        assert K == 4 + 4 + 2 + 1
        assert np.array_equal(idx.cpu().numpy(), ri) and np.array_equal(dist.cpu().numpy(), rd)


def test_bbknn_duplicates_and_distances_csr():
    """Coincident points across and within batches: zero distances are dropped from the
    distances CSR, the cell itself stays first in its own batch."""
    from metalsinglecell import _hip
    from oracle import bbknn
    X, _ = _mixture(900, 16, 3)
    X[300:330] = X[0:30]      # copies in another batch
    X[600:610] = X[610:620]   # copies within one batch
    batch = np.repeat([0, 1, 2], 300)
    idx, dist = _hip.bbknn(X, batch, 3, 5, precision="bf16")
    ri, rd, K = bbknn.bbknn_knn(X, batch, 5)
    assert np.array_equal(idx.cpu().numpy(), ri) and np.array_equal(dist.cpu().numpy(), rd)
    ip, ind, dat = _hip.bbknn_distances_csr(idx, dist)
    D = sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(900, 900))
    assert _same_csr(D, bbknn.distances_csr(ri, rd, 900))


@pytest.mark.parametrize("trim", [None, 20, 3])
def test_bbknn_graph_matches_oracle(trim):
    from metalsinglecell.neighbors import bbknn as hip_bbknn
    from oracle import bbknn
    X, _ = _mixture(1800, 24, 11)
    batch = np.random.default_rng(11).choice(["p", "q", "r", "s"], 1800)
    D, C, K = hip_bbknn(X, batch, neighbors_within_batch=3, n_pcs=None, trim=trim)
    RD, RC, RK = bbknn.bbknn(X, batch, 3, trim=trim)
    assert K == RK == 12
    assert _same_csr(D, RD)
    assert _same_csr(C, RC)


@pytest.mark.parametrize("nb,k", [(80, 3), (3, 30), (40, 20)])
def test_bbknn_wide_rows_match_oracle(nb, k):
    """More than 64 combined neighbours per row (K = sum_t min(k, n_t): 80 batches x 3, 3 x 30,
    40 x 20 = 800) and more than 64 batches: the combined rows, the distances CSR and the
    trimmed connectivities equal the oracle's (the reference loops over any number of
    batches, neighbors.py:491-541)."""
    from metalsinglecell import _hip
    from metalsinglecell.neighbors import bbknn as hip_bbknn
    from oracle import bbknn
    n = 2400 if nb < 40 else 3000
    X, _ = _mixture(n, 12, nb + k)
    batch = np.random.default_rng(nb * k).integers(0, nb, n)
    codes, m = bbknn.batch_codes(batch)
    idx, dist = _hip.bbknn(X, codes, m, k, precision="bf16")
    ri, rd, K = bbknn.bbknn_knn(X, batch, k)
    assert K > 64 and idx.shape[1] == K
    assert np.array_equal(idx.cpu().numpy(), ri) and np.array_equal(dist.cpu().numpy(), rd)
    ip, ind, dat = _hip.bbknn_distances_csr(idx, dist)
    D = sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))
    assert _same_csr(D, bbknn.distances_csr(ri, rd, n))
    HD, HC, HK = hip_bbknn(X, batch, neighbors_within_batch=k, n_pcs=None)
    RD, RC, RK = bbknn.bbknn(X, batch, k)
    assert HK == RK == K
    assert _same_csr(HD, RD) and _same_csr(HC, RC)


def test_bbknn_width_limit():
    from metalsinglecell import _hip
    X, _ = _mixture(1100, 8, 2)
    batch = np.arange(1100) % 1025
    with pytest.raises(RuntimeError, match="n_batches"):
        _hip.bbknn(X, batch, 1025, 1)


def test_trim_graph_hubs():
    """Rows far longer than trim (hubs) and tied values: thresholds by the bitwise search."""
    from metalsinglecell import _hip
    from oracle import bbknn
    rng = np.random.default_rng(4)
    n = 3000
    rows = np.r_[np.zeros(2500, int), np.ones(700, int), rng.integers(0, n, 20000)]
    cols = np.r_[rng.choice(np.arange(2, n), 2500, replace=False), rng.choice(np.arange(2, n), 700, replace=False),
                 rng.integers(0, n, 20000)]
    vals = np.round(rng.uniform(0.01, 1.0, rows.size), 2).astype(np.float32)
    A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    A = sp.csr_matrix(A.maximum(A.T), dtype=np.float32)
    A.setdiag(0)
    A.eliminate_zeros()
    A.sort_indices()
    for trim in (1, 7, 64, 1000):
        ip, ind, dat = _hip.trim_graph(A.indptr, A.indices, A.data, n, trim)
        T = sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))
        assert _same_csr(T, bbknn.trim_graph(A, trim)), trim


def test_pp_bbknn_writes_slots():
    from metalsinglecell import pp
    from metalsinglecell._anndata import AnnData
    from oracle import bbknn
    import pandas as pd
    X, _ = _mixture(1000, 40, 2)
    a = AnnData(X=X.copy())
    a.obsm["X_pca"] = X.copy()
    a.obs["batch"] = pd.Categorical(np.random.default_rng(2).choice(["b1", "b2"], 1000))
    pp.bbknn(a, batch_key="batch", n_pcs=30)
    RD, RC, K = bbknn.bbknn(X[:, :30], a.obs["batch"].to_numpy(), 3)
    assert _same_csr(a.obsp["distances"], RD) and _same_csr(a.obsp["connectivities"], RC)
    p = a.uns["neighbors"]["params"]
    assert p["n_neighbors"] == K == 6 and p["bbknn"]["trim"] == 60 and p["n_pcs"] == 30
