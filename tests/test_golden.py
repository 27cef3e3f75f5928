"""Golden fixtures (tests/golden, written by make_golden.py from the oracle): the oracle
reproduces them on the CPU and the HIP path reproduces them on the GPU."""

import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT

G = os.path.join(ROOT, "tests", "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _P(f):
    n = f["indptr"].size - 1
    return sp.csr_matrix((f["data"], f["indices"].astype(np.int32), f["indptr"]), shape=(n, n))


def test_oracle_reproduces_fixtures():
    from oracle import fuzzy, grid, knn, leiden, moran
    k = _load("knn_mixture600_k15.npz")
    idx, d2, dist = knn.exact_knn(k["X"], 15)
    assert np.array_equal(idx, k["idx"]) and np.array_equal(d2, k["d2"])
    f = _load("fuzzy_mixture600_k15.npz")
    P, sigma, rho = fuzzy.fuzzy_simplicial_set(idx, dist)
    assert np.array_equal(P.data, f["data"]) and np.array_equal(sigma, f["sigma"])
    lab = leiden.leiden(P, n_iterations=2)
    assert np.array_equal(lab, _load("leiden_mixture600.npz")["labels"])
    lv = _load("louvain_mixture600.npz")
    lab = leiden.louvain(P)
    assert np.array_equal(lab, lv["labels"]) and leiden.modularity_of_labels(P, lab) == float(lv["q"])
    s = _load("spatial_hex30_k6.npz")
    si, sd = grid.spatial_knn(s["xy"].astype(np.float32), 6)
    assert np.array_equal(si, s["idx"]) and np.array_equal(sd, s["dist"])
    m = _load("moran_hex25.npz")
    A, _ = grid.knn_graph(m["xy"], 6)
    I, sims = moran.moran_dense(m["X"], moran.row_normalize(A), moran.permutations(625, 99, 0))
    np.testing.assert_allclose(I, m["I"], atol=1e-12)
    np.testing.assert_allclose(sims, m["sims"], atol=1e-12)


@pytest.mark.gpu
def test_hip_reproduces_fixtures():
    from metalsinglecell import _hip
    from oracle import grid, moran
    k = _load("knn_mixture600_k15.npz")
    for prec in ("fp32", "bf16"):
        idx, dist = _hip.knn_exact(k["X"], 15, precision=prec)
        assert np.array_equal(idx.cpu().numpy(), k["idx"])
        assert np.array_equal(dist.cpu().numpy(), np.sqrt(k["d2"]).astype(np.float32))
    f = _load("fuzzy_mixture600_k15.npz")
    (ip, ind, dat), (sig, rho) = _hip.fuzzy_connectivities(idx, dist)
    assert np.array_equal(ip.cpu().numpy(), f["indptr"])
    assert np.array_equal(ind.cpu().numpy(), f["indices"]) and np.array_equal(dat.cpu().numpy(), f["data"])
    lab, q, _ = _hip.leiden(ip, ind, dat, 600, n_iterations=2)
    L = _load("leiden_mixture600.npz")
    assert np.array_equal(lab.cpu().numpy(), L["labels"]) and q == float(L["q"])
    lab, q, _ = _hip.louvain(ip, ind, dat, 600)
    L = _load("louvain_mixture600.npz")
    assert np.array_equal(lab.cpu().numpy(), L["labels"]) and q == float(L["q"])
    s = _load("spatial_hex30_k6.npz")
    si, sd = _hip.spatial_knn(s["xy"].astype(np.float32), 6)
    assert np.array_equal(si.cpu().numpy(), s["idx"]) and np.array_equal(sd.cpu().numpy(), s["dist"])
    m = _load("moran_hex25.npz")
    A, _ = grid.knn_graph(m["xy"], 6)
    I, sims, _ = _hip.moran(m["X"], A.indptr, A.indices, A.data.astype(np.float64), n_perms=99,
                            perms=moran.permutations(625, 99, 0), transformation=True)
    np.testing.assert_allclose(I.cpu().numpy(), m["I"], atol=1e-6, rtol=0)
    np.testing.assert_allclose(sims.cpu().numpy(), m["sims"], atol=1e-6, rtol=0)
