"""HIP Louvain (msc_louvain) against the CPU oracle: identical labels and modularity."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _canon(P):
    P = sp.csr_matrix(P, dtype=np.float32)
    P.sum_duplicates()
    P.sort_indices()
    return P


def _check(P, **kw):
    from metalsinglecell import _hip
    from oracle import leiden
    P = _canon(P)
    lab, q, info = _hip.louvain(P.indptr, P.indices, P.data, P.shape[0], **kw)
    lab = lab.cpu().numpy().astype(np.int64)
    stats = []
    ref = leiden.louvain(P, resolution=kw.get("resolution", 1.0), random_state=kw.get("random_state", 0),
                         max_levels=kw.get("max_levels", 20), commit_prob=kw.get("commit_prob", 0.9), stats=stats)
    assert np.array_equal(lab, ref), (info["levels"], stats)
    assert q == leiden.modularity_of_labels(P, ref, kw.get("resolution", 1.0))
    assert [lv["n_comms"] for lv in info["levels"]] == [s["n_comms"] for s in stats]
    assert [lv["move_passes"] for lv in info["levels"]] == [s["move_passes"] for s in stats]
    return lab, q


def _knn_graph(X, k):
    from metalsinglecell import _hip
    idx, dist = _hip.knn_exact(X, k)
    (ip, ind, dat), _ = _hip.fuzzy_connectivities(idx, dist)
    n = X.shape[0]
    return sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))


def test_louvain_mixture_3000(c1_data):
    from sklearn.metrics import adjusted_rand_score
    X, lab_true = c1_data
    lab, _ = _check(_knn_graph(X, 15))
    assert adjusted_rand_score(lab_true, lab) > 0.95


@pytest.mark.parametrize("res,seed", [(0.5, 1), (2.0, 2), (8.0, 3)])
def test_louvain_resolutions(c1_data, res, seed):
    _check(_knn_graph(c1_data[0][:1500], 15), resolution=res, random_state=seed)


def test_louvain_limits_and_max_levels(c1_data):
    """resolution 1e6: nobody moves (identity labels, one level record); max_levels=1 and
    commit_prob=1 follow the oracle too."""
    P = _knn_graph(c1_data[0][:1000], 10)
    lab, _ = _check(P, resolution=1e6)
    assert np.array_equal(lab, np.arange(1000))
    _check(P, max_levels=1)
    _check(P, commit_prob=1.0, random_state=5)


def test_louvain_heavy_vertices():
    """Hubs in every move tier (degrees 90, 300, 1500 and > 4096)."""
    rng = np.random.default_rng(6)
    n = 8000
    rows, cols = [], []
    for hub, deg in [(0, 5000), (1, 1500), (2, 300), (3, 90)]:
        nb = rng.choice(np.arange(4, n), deg, replace=False)
        rows += [hub] * deg
        cols += nb.tolist()
    for b in range(n // 100):
        m = np.arange(b * 100, (b + 1) * 100)
        rows += rng.choice(m, 800).tolist()
        cols += rng.choice(m, 800).tolist()
    w = rng.uniform(0.1, 1.0, len(rows)).astype(np.float32)
    A = sp.coo_matrix((w, (rows, cols)), shape=(n, n)).tocsr()
    A = (A + A.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    assert np.diff(A.indptr).max() > 4096
    _check(A, random_state=2)


def test_tl_louvain_writes_obs_and_params(c1_data):
    from metalsinglecell import pp, tl
    from metalsinglecell._anndata import AnnData
    from oracle import leiden
    X = c1_data[0][:1200]
    a = AnnData(X=X.copy())
    a.obsm["X_pca"] = X.copy()
    pp.neighbors(a, n_neighbors=15, use_rep="X_pca")
    tl.louvain(a, backend="gpu", random_state=4, key_added="lv")
    ref = leiden.louvain(a.obsp["connectivities"], random_state=4)
    assert np.array_equal(a.obs["lv"].astype(str).to_numpy(), np.array([str(v) for v in ref]))
    assert a.uns["lv"]["params"] == {"resolution": 1.0, "random_state": 4}
