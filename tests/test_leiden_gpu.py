"""HIP Leiden against the CPU oracle: identical labels and modularity."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _gpu_leiden(P, **kw):
    from metalsinglecell import _hip
    P = sp.csr_matrix(P, dtype=np.float32)
    P.sum_duplicates()
    P.sort_indices()
    lab, q, info = _hip.leiden(P.indptr, P.indices, P.data, P.shape[0], **kw)
    return lab.cpu().numpy().astype(np.int64), q, info


def _check(P, **kw):
    from oracle import leiden
    lab, q, info = _gpu_leiden(P, **kw)
    stats = []
    ref = leiden.leiden(P, resolution=kw.get("resolution", 1.0),
                        random_state=kw.get("random_state", 0),
                        n_iterations=kw.get("n_iterations", 2), stats=stats)
    rq = leiden.modularity_of_labels(P, ref, kw.get("resolution", 1.0))
    assert np.array_equal(lab, ref), (info["levels"], stats)
    assert q == rq
    return lab, q, info


def _knn_graph(X, k):
    from metalsinglecell import _hip
    idx, dist = _hip.knn_exact(X, k)
    (ip, ind, dat), _ = _hip.fuzzy_connectivities(idx, dist)
    n = X.shape[0]
    return sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))


def test_leiden_mixture_3000(c1_data):
    from sklearn.metrics import adjusted_rand_score
    X, lab_true = c1_data
    P = _knn_graph(X, 15)
    lab, q, _ = _check(P)
    assert adjusted_rand_score(lab_true, lab) > 0.95


@pytest.mark.parametrize("res", [0.5, 2.0, 8.0])
def test_leiden_resolutions(c1_data, res):
    P = _knn_graph(c1_data[0][:1500], 15)
    _check(P, resolution=res, random_state=3)


def test_leiden_heavy_vertices():
    """Hubs of degree 90 (two neighbours per lane on the one-wave path), 300, 700, 1500 and
    3000 (the four block classes: 512, 1024, 2048 and 4096 sorted keys, 2 to 16 per thread),
    5000 (the 1024-thread class: 16384 keys) and 17000 (the sort path over heavy edges); the
    frontier rule applies with them present."""
    rng = np.random.default_rng(5)
    n = 20000
    rows, cols = [], []
    hubs = [(0, 17000), (1, 5000), (2, 3000), (3, 1500), (4, 700), (5, 300), (6, 90)]
    for hub, deg in hubs:
        nb = rng.choice(np.arange(len(hubs), n), deg, replace=False)
        rows += [hub] * deg
        cols += nb.tolist()
    # planted blocks of 100
    for b in range(n // 100):
        m = np.arange(b * 100, (b + 1) * 100)
        r = rng.choice(m, 800)
        c = rng.choice(m, 800)
        rows += r.tolist()
        cols += c.tolist()
    w = rng.uniform(0.1, 1.0, len(rows)).astype(np.float32)
    A = sp.coo_matrix((w, (rows, cols)), shape=(n, n)).tocsr()
    A = (A + A.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    assert np.diff(A.indptr).max() > 16384
    _check(A, random_state=1)


def test_leiden_init_labels_and_single_iteration(c1_data):
    from metalsinglecell import _hip
    from oracle import leiden
    P = _knn_graph(c1_data[0][:2000], 15)
    lab1, _, _ = _gpu_leiden(P, n_iterations=1)
    ref1 = leiden.leiden(P, n_iterations=1)
    assert np.array_equal(lab1, ref1)


def test_large_graph_70k():
    """70,000 cells, 12 clusters, k = 15 (1M-edge graph): Leiden and Louvain labels and Q equal
    the oracle's."""
    from metalsinglecell import _hip, synthetic
    from oracle import leiden
    X, _ = synthetic.gaussian_mixture_isotropic(n=70_000, d=20, n_clusters=12, seed=3)
    P = _knn_graph(X, 15)
    assert P.shape[0] >= 1 << 16
    _, _, info = _check(P)
    # sigma_tot is built once per phase, then kept current from the committed vertices
    top = info["levels"][0]
    assert top["move_delta_passes"] == top["move_passes"] - 1 > 0
    assert top["refine_delta_passes"] == top["refine_passes"] - 1 > 0
    lab, q, linfo = _hip.louvain(P.indptr, P.indices, P.data, P.shape[0])
    ref = leiden.louvain(P)
    assert np.array_equal(lab.cpu().numpy().astype(np.int64), ref)
    assert q == leiden.modularity_of_labels(P, ref)
    assert linfo["levels"][0]["move_delta_passes"] > 0


def _planted_partition(n, group, k_in, k_out, seed):
    """Groups of `group` vertices, k_in random in-group and k_out random out-group edges per
    vertex (symmetrised, uniform weights), vertices shuffled as real cell orders are."""
    rng = np.random.default_rng(seed)
    g = np.arange(n) // group
    rows = np.repeat(np.arange(n), k_in + k_out)
    cols = np.empty((n, k_in + k_out), np.int64)
    cols[:, :k_in] = (g * group)[:, None] + rng.integers(0, group, (n, k_in))
    cols[:, k_in:] = rng.integers(0, n, (n, k_out))
    w = rng.uniform(0.2, 1.0, rows.size).astype(np.float32)
    A = sp.coo_matrix((w, (rows, cols.ravel())), shape=(n, n)).tocsr()
    A = (A + A.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    perm = rng.permutation(n)
    A = A[perm][:, perm].tocsr()
    A.sort_indices()
    return A.astype(np.float32)


def test_large_graph_planted_300k():
    """300,000 shuffled vertices in 3,000 planted groups (4M edges): Leiden and Louvain labels and
    Q equal the oracle's, with sigma_tot kept from the committed vertices after each phase's first pass."""
    from metalsinglecell import _hip
    from oracle import leiden
    P = _planted_partition(300_000, 100, 6, 1, 11)
    _, _, info = _check(P)
    assert info["levels"][0]["move_delta_passes"] > 0 and info["levels"][0]["refine_delta_passes"] > 0
    lab, q, _ = _hip.louvain(P.indptr, P.indices, P.data, P.shape[0])
    ref = leiden.louvain(P)
    assert np.array_equal(lab.cpu().numpy().astype(np.int64), ref)
    assert q == leiden.modularity_of_labels(P, ref)
