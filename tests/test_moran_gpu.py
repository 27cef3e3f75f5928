"""HIP Moran's I against the float64 oracle (1e-6 bar) and the reference's known answers."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _lattice_signals(side=25, jitter=0.0, seed=0):
    """tests/test_spatial_autocorr.py:24-42 of the reference: smooth, checker, noise genes."""
    r, c = np.divmod(np.arange(side * side), side)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6]
    if jitter:
        xy = xy + np.random.default_rng(seed).normal(0, jitter, xy.shape)
    rng = np.random.default_rng(seed)
    X = np.c_[(r + c).astype(np.float32), ((r + c) % 2).astype(np.float32),
              rng.normal(size=side * side)].astype(np.float32)
    return xy, X


def _knn_W(xy, k=6):
    from oracle import grid
    A, _ = grid.knn_graph(xy, k)
    return A


def _gpu(X, W, n_perms=0, perms=None, seed=0, transformation=True):
    from metalsinglecell import _hip
    W = sp.csr_matrix(W)
    I, sims, S0 = _hip.moran(X, W.indptr, W.indices, W.data.astype(np.float64), n_perms=n_perms,
                             perms=perms, seed=seed, transformation=transformation)
    return I.cpu().numpy(), (None if sims is None else sims.cpu().numpy()), S0


@pytest.mark.parametrize("transformation", [True, False])
def test_moran_matches_float64_oracle(transformation):
    from oracle import moran
    xy, X = _lattice_signals(22, jitter=8.0, seed=1)
    W = _knn_W(xy)
    Wn = moran.row_normalize(W) if transformation else sp.csr_matrix(W, dtype=np.float64)
    perms = moran.permutations(X.shape[0], 99, 0)
    I, sims, _ = _gpu(X, W, 99, perms, transformation=transformation)
    rI, rs = moran.moran_dense(X, Wn, perms)
    np.testing.assert_allclose(I, rI, atol=1e-6, rtol=0)
    np.testing.assert_allclose(sims, rs, atol=1e-6, rtol=0)
    sI, ss = moran.moran_scatter(X, Wn, perms)
    np.testing.assert_allclose(I, sI, atol=1e-6, rtol=0)


def test_moran_known_answers_999_perms():
    """Reference KAT tests/test_spatial_autocorr.py:61-70 (checker I<0 and smooth I>0, both
    at pval_sim = 1/1000 with 999 permutations; noise > 0.05)."""
    from oracle import moran
    xy, X = _lattice_signals(25)
    W = _knn_W(xy)
    perms = moran.permutations(X.shape[0], 999, 0)
    I, sims, _ = _gpu(X, W, 999, perms)
    p_sim, _, _ = moran.permutation_pvals(I, sims)
    assert I[1] < 0 < I[0]
    assert abs(p_sim[0] - 1 / 1000) < 1e-9 and abs(p_sim[1] - 1 / 1000) < 1e-9
    assert p_sim[2] > 0.05


def test_moran_device_permutations_distribution():
    """Feistel permutations (no host stream) give the same null distribution as the
    reference's numpy permutation stream: per-gene null means and variances agree within
    sampling error, and the device null is deterministic."""
    from oracle import moran
    rng = np.random.default_rng(7)
    n = 4000
    xy = rng.uniform(0, 1000, (n, 2))
    X = rng.normal(size=(n, 40)).astype(np.float32)
    W = _knn_W(xy)
    P = 400
    I, sims_dev, _ = _gpu(X, W, P, None, seed=11)
    _, sims_np, _ = _gpu(X, W, P, moran.permutations(n, P, 0))
    v_dev, v_np = sims_dev.var(axis=0), sims_np.var(axis=0)
    assert 0.85 < v_dev.mean() / v_np.mean() < 1.15
    se = np.sqrt(v_np.mean() / P)
    assert abs(sims_dev.mean() - sims_np.mean()) < 4 * se / np.sqrt(sims_np.shape[1]) + 1e-4
    assert abs(sims_dev.mean() - (-1 / (n - 1))) < 4 * se
    I2, sims2, _ = _gpu(X, W, P, None, seed=11)
    assert np.array_equal(sims_dev, sims2)          # deterministic


def test_moran_constant_gene_is_nan():
    xy, X = _lattice_signals(10)
    X = np.c_[X, np.full(X.shape[0], 3.0, np.float32)]
    I, _, _ = _gpu(X, _knn_W(xy))
    assert np.isnan(I[-1]) and np.isfinite(I[:-1]).all()


@pytest.mark.parametrize("n_genes", [8, 7])
def test_moran_wide_genes_match_oracle(n_genes):
    """G % 4 == 0 takes the four-genes-per-thread (16-byte) gather path, other G the scalar
    one; both must meet the float64 oracle at 1e-6."""
    from oracle import moran
    xy, X3 = _lattice_signals(20, jitter=5.0, seed=2)
    rng = np.random.default_rng(3)
    X = np.c_[X3, rng.gamma(2.0, 1.0, (X3.shape[0], n_genes - 3))].astype(np.float32)
    W = _knn_W(xy)
    perms = moran.permutations(X.shape[0], 37, 5)
    I, sims, _ = _gpu(X, W, 37, perms)
    rI, rs = moran.moran_dense(X, moran.row_normalize(W), perms)
    np.testing.assert_allclose(I, rI, atol=1e-6, rtol=0)
    np.testing.assert_allclose(sims, rs, atol=1e-6, rtol=0)
