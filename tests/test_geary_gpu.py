"""HIP Geary's C (msc_geary) against the float64 edge-form oracle, and the gr frame contract."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _lattice_signals(side=25, jitter=0.0, seed=0):
    """The reference's smooth / checker / noise lattice (tests/test_spatial_autocorr.py:24-42)."""
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
    C, sims, S0 = _hip.geary(X, W.indptr, W.indices, W.data.astype(np.float64), n_perms=n_perms,
                             perms=perms, seed=seed, transformation=transformation)
    return C.cpu().numpy(), (None if sims is None else sims.cpu().numpy()), S0


@pytest.mark.parametrize("transformation", [True, False])
def test_geary_matches_float64_oracle(transformation):
    from oracle import moran
    xy, X = _lattice_signals(22, jitter=8.0, seed=1)
    W = _knn_W(xy)
    Wn = moran.row_normalize(W) if transformation else sp.csr_matrix(W, dtype=np.float64)
    perms = moran.permutations(X.shape[0], 99, 0)
    C, sims, S0 = _gpu(X, W, 99, perms, transformation=transformation)
    rC, rs = moran.geary_scatter(X, Wn, perms)
    np.testing.assert_allclose(C, rC, atol=1e-9, rtol=0)
    np.testing.assert_allclose(sims, rs, atol=1e-9, rtol=0)
    assert S0 == pytest.approx(float(Wn.sum()), rel=1e-14)


@pytest.mark.parametrize("G", [8, 7])
def test_geary_many_genes_both_gather_widths(G):
    """G % 4 == 0 takes the 16-byte gather path, G = 7 the scalar one; asymmetric weights."""
    from oracle import moran
    rng = np.random.default_rng(G)
    n = 1500
    X = rng.normal(size=(n, G)).astype(np.float32)
    X[:, 0] = np.linspace(0, 5, n)
    A = sp.random(n, n, density=0.004, random_state=rng, format="csr")
    A.data = rng.uniform(0.1, 2.0, A.data.size)
    perms = moran.permutations(n, 17, 5)
    C, sims, _ = _gpu(X, A, 17, perms, transformation=False)
    rC, rs = moran.geary_scatter(X, A, perms)
    np.testing.assert_allclose(C, rC, atol=1e-9, rtol=0)
    np.testing.assert_allclose(sims, rs, atol=1e-9, rtol=0)


def test_geary_device_permutations_center_on_one():
    """The device Feistel null: E[C] = 1 under randomisation; the checkerboard (C > 1) and the
    smooth gene (C < 1) sit in opposite tails."""
    xy, X = _lattice_signals(25)
    C, sims, _ = _gpu(X, _knn_W(xy), 400, None, seed=3)
    assert abs(sims[:, 2].mean() - 1.0) < 0.02
    assert C[0] < sims[:, 0].min() and C[1] > sims[:, 1].max()


@pytest.mark.parametrize("mode,stat", [("moran", "I"), ("geary", "C")])
def test_gr_frame_columns_and_sort_order(mode, stat):
    """Reference tests/test_spatial_autocorr.py:73-88 (columns, ascending C / descending I)."""
    from metalsinglecell import gr
    from metalsinglecell._anndata import AnnData
    import pandas as pd
    xy, X = _lattice_signals(25)
    a = AnnData(X=X.copy(), var=pd.DataFrame(index=["smooth", "checker", "noise"]))
    a.obsm["spatial"] = xy
    gr.spatial_neighbors_knn(a, n_neighs=6)
    gr.spatial_autocorr(a, mode=mode)
    df = a.uns["moranI" if mode == "moran" else "gearyC"]
    assert df.columns.tolist() == [stat, "pval_norm", "var_norm", "pval_norm_fdr_bh"]
    gr.spatial_autocorr(a, mode=mode, n_perms=99)
    df = a.uns["moranI" if mode == "moran" else "gearyC"]
    assert df.columns.tolist() == [stat, "pval_norm", "var_norm", "pval_z_sim", "pval_sim", "var_sim",
                                   "pval_norm_fdr_bh", "pval_z_sim_fdr_bh", "pval_sim_fdr_bh"]
    order = df[stat].to_numpy()
    assert np.all(np.diff(order) <= 0) if mode == "moran" else np.all(np.diff(order) >= 0)
