"""AnnData-level contract of the neighbour-graph path on the device: ``pp.neighbors`` ->
``tl.leiden(backend="gpu")`` / ``tl.louvain(backend="gpu")`` and ``cluster.leiden``, every slot
compared with the oracle.

Mirrors the reference's contract tests: the ``pp.neighbors`` slots and recall
(tests/test_gpu_parity.py:80-95), natsorted categories at 11+ clusters
(tests/test_tl_contracts.py:148-157), recorded params (:182-187) and ``neighbors_key`` /
``key_added`` routing (:276-286).
"""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _adata(n=2400, d=50, n_clusters=24, seed=7, n_vars=80):
    """A wide object (n_vars > 50) whose representation is a planted 24-cluster X_pca."""
    from metalsinglecell import synthetic
    from metalsinglecell._anndata import AnnData
    Xp, lab = synthetic.gaussian_mixture_isotropic(n=n, d=d, n_clusters=n_clusters, seed=seed)
    a = AnnData(np.zeros((n, n_vars), np.float32))
    a.obsm["X_pca"] = Xp
    return a, lab


def _oracle_graph(X, k):
    from oracle import fuzzy, knn
    ri, _, rd = knn.exact_knn_c(X, k)
    D = knn.distances_csr(ri, rd, X.shape[0])
    C, _, _ = fuzzy.fuzzy_simplicial_set(ri, rd)
    return D, C


def _equal(A, B):
    A, B = sp.csr_matrix(A), sp.csr_matrix(B)
    A.sort_indices()
    B.sort_indices()
    return (A.shape == B.shape and np.array_equal(A.indptr, B.indptr)
            and np.array_equal(A.indices, B.indices) and np.array_equal(A.data, B.data)
            and A.data.dtype == B.data.dtype)


def test_pp_neighbors_slots_equal_the_oracle():
    from metalsinglecell import pp
    a, _ = _adata()
    assert pp.neighbors(a, 15) is None
    D, C = _oracle_graph(a.obsm["X_pca"], 15)
    assert _equal(a.obsp["distances"], D)
    assert _equal(a.obsp["connectivities"], C)
    assert a.obsp["distances"].dtype == np.float32 and a.obsp["connectivities"].dtype == np.float32
    assert np.all(np.diff(a.obsp["distances"].indptr) == 14)
    assert a.uns["neighbors"] == {
        "connectivities_key": "connectivities", "distances_key": "distances",
        "params": {"n_neighbors": 15, "method": "umap", "random_state": 0, "metric": "euclidean"}}


def test_pp_neighbors_n_pcs_truncates_and_is_recorded():
    from metalsinglecell import pp
    a, _ = _adata()
    pp.neighbors(a, 15, 20)                     # n_pcs is positional after n_neighbors
    D, C = _oracle_graph(np.ascontiguousarray(a.obsm["X_pca"][:, :20]), 15)
    assert _equal(a.obsp["distances"], D) and _equal(a.obsp["connectivities"], C)
    assert a.uns["neighbors"]["params"]["n_pcs"] == 20
    assert "use_rep" not in a.uns["neighbors"]["params"]
    pp.neighbors(a, 10, use_rep="X_pca")
    assert a.uns["neighbors"]["params"]["use_rep"] == "X_pca"
    assert a.uns["neighbors"]["params"]["n_neighbors"] == 10


def test_pp_neighbors_copy_leaves_the_input_untouched():
    from metalsinglecell import pp
    a, _ = _adata(n=600, n_clusters=6)
    b = pp.neighbors(a, 15, copy=True)
    assert "distances" not in a.obsp and "neighbors" not in a.uns
    assert "connectivities" in b.obsp and b.uns["neighbors"]["params"]["n_neighbors"] == 15


def test_key_added_and_neighbors_key_route_to_the_right_graph():
    from metalsinglecell import pp, tl
    from oracle import leiden
    a, _ = _adata()
    pp.neighbors(a, 15)
    pp.neighbors(a, 5, key_added="small")
    assert set(a.obsp) == {"distances", "connectivities", "small_distances", "small_connectivities"}
    assert a.uns["small"]["connectivities_key"] == "small_connectivities"
    assert a.uns["small"]["distances_key"] == "small_distances"
    D5, C5 = _oracle_graph(a.obsm["X_pca"], 5)
    assert _equal(a.obsp["small_distances"], D5) and _equal(a.obsp["small_connectivities"], C5)
    _, C15 = _oracle_graph(a.obsm["X_pca"], 15)
    tl.leiden(a, backend="gpu", neighbors_key="small", key_added="leiden_small")
    tl.leiden(a, backend="gpu")
    tl.louvain(a, backend="gpu", neighbors_key="small", key_added="louvain_small")
    small = leiden.leiden(C5, resolution=1.0, random_state=0, n_iterations=2)
    full = leiden.leiden(C15, resolution=1.0, random_state=0, n_iterations=2)
    lsmall = leiden.louvain(C5, resolution=1.0, random_state=0)
    assert np.array_equal(a.obs["leiden_small"].astype(str).to_numpy(), small.astype(str))
    assert np.array_equal(a.obs["leiden"].astype(str).to_numpy(), full.astype(str))
    assert np.array_equal(a.obs["louvain_small"].astype(str).to_numpy(), lsmall.astype(str))
    assert not np.array_equal(small, full)      # the two graphs really differ


def test_leiden_categories_natsorted_and_params_recorded():
    from metalsinglecell import pp, tl
    from oracle import leiden
    a, planted = _adata()
    pp.neighbors(a, 15)
    tl.leiden(a, backend="gpu", random_state=3, key_added="fine")
    cats = list(a.obs["fine"].cat.categories)
    assert len(cats) >= 11, f"only {len(cats)} clusters"
    assert cats == sorted(cats, key=int) and cats != sorted(cats)
    codes = a.obs["fine"].cat.codes.to_numpy()
    labels = a.obs["fine"].astype(str).astype(int).to_numpy()
    assert np.array_equal(codes, labels)        # natsort keeps code == integer label
    _, C = _oracle_graph(a.obsm["X_pca"], 15)
    assert np.array_equal(labels, leiden.leiden(C, resolution=1.0, random_state=3, n_iterations=2))
    assert a.uns["fine"] == {"params": {"resolution": 1.0, "random_state": 3, "n_iterations": 2}}
    tl.louvain(a, backend="gpu", random_state=3)
    assert a.uns["louvain"] == {"params": {"resolution": 1.0, "random_state": 3}}
    from sklearn.metrics import adjusted_rand_score
    assert adjusted_rand_score(planted, labels) >= 0.95


def test_cluster_leiden_gpu_branch_equals_oracle():
    from metalsinglecell import cluster
    from oracle import leiden
    a, _ = _adata(n=1500, n_clusters=12, seed=11)
    _, C = _oracle_graph(a.obsm["X_pca"], 15)
    for res, it in ((1.0, 2), (0.5, 1), (2.0, 3)):
        got = cluster.leiden(C, resolution=res, random_state=5, n_iterations=it, backend="gpu")
        assert got.dtype == np.int64
        assert np.array_equal(got, leiden.leiden(C, resolution=res, random_state=5, n_iterations=it))
    coo = C.tocoo()                              # a non-canonical input is canonicalised first
    perm = np.random.default_rng(0).permutation(coo.nnz)
    shuffled = sp.coo_matrix((coo.data[perm], (coo.row[perm], coo.col[perm])), shape=C.shape)
    assert np.array_equal(cluster.leiden(shuffled, backend="gpu"), leiden.leiden(C))
