"""HIP fuzzy-simplicial-set connectivities against the CPU oracle."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _gpu_conn(X, k):
    from metalsinglecell import _hip
    idx, dist = _hip.knn_exact(X, k)
    (ip, ind, dat), (sig, rho) = _hip.fuzzy_connectivities(idx, dist)
    n = X.shape[0]
    P = sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))
    return idx.cpu().numpy(), dist.cpu().numpy(), P, sig.cpu().numpy(), rho.cpu().numpy()


def _check(X, k):
    from oracle import fuzzy
    idx, dist, P, sig, rho = _gpu_conn(X, k)
    R, rsig, rrho = fuzzy.fuzzy_simplicial_set(idx, dist)
    assert np.array_equal(rho, rrho)
    assert np.array_equal(sig, rsig)
    assert np.array_equal(P.indptr, R.indptr) and np.array_equal(P.indices, R.indices)
    assert np.array_equal(P.data, R.data)          # bit-identical (tolerance 1e-5 is the spec)
    assert P.has_canonical_format
    assert (abs(P - P.T) > 0).nnz == 0


def test_fuzzy_mixture_3000(c1_data):
    _check(c1_data[0], 15)


@pytest.mark.parametrize("k", [2, 5, 30])
def test_fuzzy_gaussian_k(k):
    X = np.random.default_rng(1).standard_normal((700, 12)).astype(np.float32)
    _check(X, k)


def test_fuzzy_duplicates_no_nan():
    """Reference tests/test_postfix_fixes.py:197-207: duplicate points give no NaN."""
    rng = np.random.default_rng(0)
    X = rng.standard_normal((60, 10)).astype(np.float32)
    Xd = np.vstack([X, X[:5], X[:5], X[:5]]).astype(np.float32)
    _check(Xd, 10)
    _, _, P, _, _ = _gpu_conn(Xd, 10)
    assert not np.isnan(P.data).any()


def test_fuzzy_hub_in_degree():
    """A hub (one point near everything) has in-degree far above k: exercises long in-lists."""
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3000, 3)).astype(np.float32) * 10
    X[:200] = rng.standard_normal((200, 3)).astype(np.float32) * 0.01
    _check(X, 15)


def test_fuzzy_floor_and_duplicate_branches_on_hand_built_rows():
    """The sigma floor with the row mean (ties at rho), the floor with the global mean (a row
    of duplicates, rho = 0) and rho skipping zero-distance duplicates, on a hand-built kNN:
    sigma, rho and the CSR bit-identical to the oracle, whose values these rows pin."""
    import torch
    from metalsinglecell import _hip
    from oracle import fuzzy
    k, n = 9, 24
    pattern = np.array([[0, 2, 2, 2, 2, 5, 6, 7, 8],           # floor binds, row mean
                        [0, 0, 0, 0, 0, 0, 0, 0, 0],           # rho = 0, global mean
                        [0, 0, 0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0],  # rho skips duplicates
                        [0, 1, 1.5, 2, 2.5, 3, 3.5, 4, 4.5]], np.float32)
    dist = pattern[np.arange(n) % 4]
    idx = (np.arange(n)[:, None] + np.arange(k)[None, :] * 5) % n
    assert all(np.unique(r).size == k for r in idx)
    dev = _hip.device()
    ti = torch.from_numpy(idx.astype(np.int32)).to(dev)
    td = torch.from_numpy(dist).to(dev)
    (ip, ind, dat), (sig, rho) = _hip.fuzzy_connectivities(ti, td)
    R, rsig, rrho = fuzzy.fuzzy_simplicial_set(idx, dist)
    assert np.array_equal(rho.cpu().numpy(), rrho) and np.array_equal(sig.cpu().numpy(), rsig)
    assert rrho[0] == 2.0 and rrho[1] == 0.0 and rrho[2] == np.float32(1.5)
    assert rsig[0] == np.float32(1e-3 * 34 / 9)
    assert rsig[1] == np.float32(1e-3 * float(dist.astype(np.float64).sum()) / dist.size)
    P = sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))
    assert np.array_equal(P.indptr, R.indptr) and np.array_equal(P.indices, R.indices)
    assert np.array_equal(P.data, R.data)
