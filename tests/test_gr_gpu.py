"""``gr`` on the device against the oracle: the spatial kNN / radius graph slots with
``library_key``, ``percentile``, ``set_diag`` and ``transform`` (dense per-section oracle), and
the ``uns['moranI']`` / ``uns['gearyC']`` frames value by value (float64 Moran / Geary oracle,
pysal variance, folded permutation p-values, Benjamini-Hochberg by its definition).

Reference contracts: tests/test_spatial_neighbors.py:79-85 (fixed degree, binary data),
:249-294 (per-library blocks), tests/test_spatial_autocorr.py:61-143 (frame columns, order,
p-values)."""

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _tissue(n=1500, seed=3, n_libs=3):
    """Interleaved libraries whose coordinate frames overlap (a graph across them would link
    cells of different sections), with some duplicated spots."""
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, n_libs, n)
    xy = rng.uniform(0, 2000, (n, 2)) + codes[:, None] * 300.0
    dup = np.flatnonzero(codes == 0)[:10]
    xy[dup[5:]] = xy[dup[:5]]
    from metalsinglecell._anndata import AnnData
    a = AnnData(np.zeros((n, 3), np.float32))
    a.obsm["spatial"] = xy
    a.obs["lib"] = pd.Categorical([f"slide{c}" for c in codes])
    return a, xy, codes


@pytest.mark.parametrize("library", [False, True])
@pytest.mark.parametrize("percentile,set_diag,transform", [
    (None, False, None), (80.0, True, None), (None, True, "spectral"), (50.0, False, "cosine")])
def test_spatial_neighbors_knn_slots(library, percentile, set_diag, transform):
    from metalsinglecell import gr
    from oracle import grid
    a, xy, codes = _tissue()
    gr.spatial_neighbors_knn(a, n_neighs=6, library_key="lib" if library else None,
                             percentile=percentile, set_diag=set_diag, transform=transform)
    A, D = grid.section_graph_dense(xy, codes if library else None, n_neighs=6,
                                    percentile=percentile, set_diag=set_diag, transform=transform)
    adj, dist = a.obsp["spatial_connectivities"], a.obsp["spatial_distances"]
    assert np.array_equal(dist.toarray(), D)
    assert dist.dtype == np.float64 and not (dist.data == 0).any()
    if transform is None:
        assert adj.dtype == np.float32 and np.array_equal(adj.toarray(), A)
        if percentile is None:
            assert np.all(np.diff(adj.indptr) == 6 + int(set_diag))
    else:
        tol = 1e-12 if transform == "spectral" else 1e-6
        np.testing.assert_allclose(adj.toarray(), A, rtol=tol, atol=tol * 1e-3)
    assert a.uns["spatial_neighbors"] == {
        "connectivities_key": "spatial_connectivities", "distances_key": "spatial_distances",
        "params": {"coord_type": "generic", "n_neighbors": 6, "transform": transform}}


def test_spatial_neighbors_radius_with_libraries_and_key_added():
    from metalsinglecell import gr
    from oracle import grid
    a, xy, codes = _tissue(n=1200, seed=5)
    gr.spatial_neighbors_radius(a, radius=(10.0, 90.0), library_key="lib", percentile=70.0,
                                set_diag=True, key_added="ring")
    A, D = grid.section_graph_dense(xy, codes, radius=(10.0, 90.0), percentile=70.0, set_diag=True)
    assert np.array_equal(a.obsp["ring_connectivities"].toarray(), A)
    assert np.array_equal(a.obsp["ring_distances"].toarray(), D)
    assert a.uns["ring_neighbors"]["params"] == {"coord_type": "generic", "radius": (10.0, 90.0),
                                                 "transform": None}
    res = gr.spatial_neighbors_radius(a, radius=90.0, library_key="lib", copy=True)
    A2, D2 = grid.section_graph_dense(xy, codes, radius=90.0)
    assert np.array_equal(res.connectivities.toarray(), A2)
    assert np.array_equal(res.distances.toarray(), D2)


def _expression(n, n_genes=24, seed=9):
    from metalsinglecell import synthetic
    counts = synthetic.nb_counts(n, n_genes=n_genes, density=0.2, seed=seed)
    return sp.csr_matrix(counts, dtype=np.float32)


@pytest.mark.parametrize("mode", ["moran", "geary"])
def test_spatial_autocorr_frame_values(mode):
    from metalsinglecell import gr
    from oracle import moran
    a, xy, codes = _tissue(n=1500, seed=4, n_libs=1)
    n = xy.shape[0]
    X = _expression(n)
    # a smooth gene and a checker-like gene so that both tails appear
    smooth = (xy[:, 0] / 200.0).astype(np.float32)
    X = sp.hstack([X, sp.csr_matrix(smooth[:, None])]).tocsr()
    a.X = X
    a.var = pd.DataFrame(index=[f"g{i}" for i in range(X.shape[1])])
    hv = np.ones(X.shape[1], bool)
    hv[[2, 5]] = False
    a.var["highly_variable"] = hv
    gr.spatial_neighbors_knn(a, n_neighs=6)
    P = 99
    gr.spatial_autocorr(a, mode=mode, n_perms=P, seed=7)
    df = a.uns["moranI" if mode == "moran" else "gearyC"]
    genes = [g for g, keep in zip(a.var_names, hv) if keep]
    assert sorted(df.index) == sorted(genes)
    W = moran.row_normalize(a.obsp["spatial_connectivities"])
    Xg = X[:, np.flatnonzero(hv)].toarray().astype(np.float64)
    perms = moran.permutations(n, P, 7)
    stat = "I" if mode == "moran" else "C"
    if mode == "moran":
        score, sims = moran.moran_scatter(Xg, W, perms)
        p_norm, var_norm = moran.analytic_pval(score, sp.csr_matrix(W))
    else:
        from scipy import stats
        score, sims = moran.geary_scatter(Xg, W, perms)
        var_norm = moran.geary_variance(sp.csr_matrix(W))
        z = (score - 1.0) / np.sqrt(var_norm)
        p_norm = np.where(z > 0, 1.0 - stats.norm.cdf(z), stats.norm.cdf(z))
    p_sim, p_z, var_sim = moran.permutation_pvals(score, sims)
    ref = pd.DataFrame({stat: score, "pval_norm": p_norm, "var_norm": np.full(score.shape, var_norm),
                        "pval_z_sim": p_z, "pval_sim": p_sim, "var_sim": var_sim}, index=genes)
    for c in ("pval_norm", "pval_z_sim", "pval_sim"):
        ref[f"{c}_fdr_bh"] = moran.bh(ref[c].to_numpy())
    assert df.columns.tolist() == ref.columns.tolist()
    got = df.loc[genes]
    # the device keeps Y = W Xc in float32 with float64 sums: I within 1e-7 (the bar is 1e-6)
    np.testing.assert_allclose(got[stat], ref[stat], rtol=0, atol=1e-7)
    np.testing.assert_allclose(got["var_norm"], ref["var_norm"], rtol=1e-12)
    np.testing.assert_allclose(got["var_sim"], ref["var_sim"], rtol=1e-6, atol=1e-15)
    for c in ("pval_norm", "pval_z_sim", "pval_norm_fdr_bh", "pval_z_sim_fdr_bh"):
        np.testing.assert_allclose(got[c], ref[c], rtol=1e-5, atol=1e-9)
    # sparse count genes put many permuted sums exactly on the observed value (in exact
    # arithmetic), so ">=" there is decided by rounding: the device count must be one the
    # oracle allows for some resolution of the sims within 1e-7 of the statistic
    folded = np.rint(got["pval_sim"].to_numpy() * (P + 1) - 1).astype(int)
    lo = (sims >= score + 1e-7).sum(axis=0)
    hi = (sims >= score - 1e-7).sum(axis=0)
    for f, a_, b_ in zip(folded, lo, hi):
        assert f in {min(c, P - c) for c in range(a_, b_ + 1)}
    exact = lo == hi
    np.testing.assert_array_equal(got["pval_sim"].to_numpy()[exact], ref["pval_sim"].to_numpy()[exact])
    vals = df[stat].to_numpy()
    assert np.all(np.diff(vals) <= 0) if mode == "moran" else np.all(np.diff(vals) >= 0)
    # the smooth gene is the most autocorrelated
    assert df.index[0] == genes[-1]
