"""Host-side logic of the pp/tl/gr layer (CPU): categories, p-values, graph tails, API errors."""

import subprocess
import sys

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp

from conftest import PKG


def test_import_does_not_load_torch_or_the_library():
    """Reference tests/test_import_contract.py:16-40: importing the package is cheap."""
    code = ("import sys; sys.path.insert(0, %r); import metalsinglecell as msc; "
            "assert 'torch' not in sys.modules; assert msc.pp and msc.tl and msc.gr; print('ok')" % PKG)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr


def test_categorical_natsorted():
    """Reference tests/test_tl_contracts.py:148-158 and tl.py:35-52."""
    from metalsinglecell.tl import _categorical
    cat = _categorical(np.array([10, 2, 1, 0, 11, 2, 3, 9]))
    assert list(cat.categories) == ["0", "1", "2", "3", "9", "10", "11"]
    assert list(cat.codes[:3]) == [5, 2, 1]


def test_moran_host_statistics_match_oracle():
    from metalsinglecell import spatial
    from oracle import moran
    rng = np.random.default_rng(0)
    W = sp.random(200, 200, density=0.05, random_state=0, format="csr")
    Wn = spatial.row_normalize(W)
    np.testing.assert_allclose(Wn.toarray(), moran.row_normalize(W).toarray(), rtol=1e-15)
    assert spatial.graph_moments(Wn) == pytest.approx(moran.graph_moments(Wn), rel=1e-12)
    score = rng.normal(0, 0.05, 20)
    for tt in (False, True):
        a = spatial.analytic_pval(score, Wn, two_tailed=tt)
        b = moran.analytic_pval(score, Wn, two_tailed=tt)
        np.testing.assert_allclose(a[0], b[0], rtol=1e-12)
        assert a[1] == pytest.approx(b[1], rel=1e-12)
    sims = rng.normal(0, 0.05, (99, 20))
    for x, y in zip(spatial.permutation_pvals(score, sims), moran.permutation_pvals(score, sims)):
        np.testing.assert_allclose(x, y, rtol=1e-12)
    for p in [rng.uniform(size=200), np.repeat(0.03, 20), np.array([1.0, 1.0, 1e-12]), np.array([0.5])]:
        np.testing.assert_allclose(spatial.multiple_testing(p), moran.bh(p), rtol=1e-12)
        np.testing.assert_allclose(spatial.multiple_testing(p, "bonferroni"), np.minimum(p * p.size, 1))
    with pytest.raises(NotImplementedError, match="corr_method"):
        spatial.multiple_testing(np.array([0.1, 0.2]), "holm")


def test_two_tailed_doubles():
    from metalsinglecell import spatial
    W = sp.random(50, 50, density=0.1, random_state=3, format="csr")
    s = np.array([0.1, -0.05, 0.2])
    one, _ = spatial.analytic_pval(s, W)
    two, _ = spatial.analytic_pval(s, W, two_tailed=True)
    np.testing.assert_allclose(two, 2 * one)


def test_moran_frame_columns_and_order():
    """Reference tests/test_spatial_autocorr.py:73-87 column contract and sort order."""
    from metalsinglecell import gr
    W = sp.random(60, 60, density=0.1, random_state=1, format="csr")
    I = np.array([0.1, -0.3, 0.5])
    df = gr.moran_frame(I, None, W, ["a", "b", "c"])
    assert df.columns.tolist() == ["I", "pval_norm", "var_norm", "pval_norm_fdr_bh"]
    assert df.index.tolist() == ["c", "a", "b"]
    sims = np.random.default_rng(0).normal(0, 0.1, (99, 3))
    df = gr.moran_frame(I, sims, W, ["a", "b", "c"])
    assert df.columns.tolist() == ["I", "pval_norm", "var_norm", "pval_z_sim", "pval_sim", "var_sim",
                                   "pval_norm_fdr_bh", "pval_z_sim_fdr_bh", "pval_sim_fdr_bh"]
    df = gr.moran_frame(I, sims, W, ["a", "b", "c"], corr_method=None)
    assert not [c for c in df.columns if "fdr" in c]


def _host_knn_edges(n_neighs):
    """A host stand-in for the device builder (the brute-force oracle), so the assembly logic
    runs on the CPU."""
    from oracle import grid

    def edges_of(xy32, xy64):
        m = xy32.shape[0]
        k = min(n_neighs, m - 1)
        idx, _ = grid.spatial_knn(xy32, k)
        rows = np.repeat(np.arange(m), k)
        lens = np.sqrt(((xy64[rows] - xy64[idx.ravel()]) ** 2).sum(1))
        return np.arange(m + 1) * k, idx.ravel(), lens
    return edges_of


@pytest.mark.parametrize("transform", [None, "spectral", "cosine"])
@pytest.mark.parametrize("set_diag", [False, True])
@pytest.mark.parametrize("percentile", [None, 60.0])
def test_section_graph_assembly_equals_dense_oracle(transform, set_diag, percentile):
    """Interleaved libraries (one with duplicate points, so zero lengths occur), per-library
    percentile, set_diag and both transforms against the dense per-section oracle."""
    import pandas as pd
    from metalsinglecell import gr
    from metalsinglecell._anndata import AnnData
    from oracle import grid
    rng = np.random.default_rng(8)
    n = 300
    codes = rng.integers(0, 3, n)
    xy = rng.uniform(0, 100, (n, 2)) + codes[:, None] * 30.0      # sections overlap in space
    dup = np.flatnonzero(codes == 1)[:6]
    xy[dup[3:]] = xy[dup[:3]]
    a = AnnData(np.zeros((n, 2), np.float32))
    a.obsm["spatial"] = xy
    a.obs["lib"] = pd.Categorical(np.array(["s0", "s1", "s2", "unused"])[codes],
                                  categories=["s2", "unused", "s0", "s1"])
    adj, dist = gr._section_graphs(a, "spatial", "lib", _host_knn_edges(6), transform,
                                   set_diag, percentile)
    A, D = grid.section_graph_dense(xy, codes, n_neighs=6, percentile=percentile,
                                    set_diag=set_diag, transform=transform)
    assert adj.shape == dist.shape == (n, n)
    assert dist.dtype == np.float64 and not (dist.data == 0).any() and dist.diagonal().sum() == 0
    assert np.array_equal(dist.toarray(), D)
    if transform is None:
        assert adj.dtype == np.float32 and np.array_equal(adj.toarray(), A)
    else:
        # sklearn's cosine keeps the float32 of the binary matrix (as the reference's call)
        tol = 1e-12 if transform == "spectral" else 1e-6
        np.testing.assert_allclose(adj.toarray(), A, rtol=tol, atol=tol * 1e-3)
    # no edge crosses between libraries
    r, c = adj.nonzero()
    assert np.array_equal(codes[r], codes[c])
    if percentile is None and transform is None:
        # each row lists its neighbours nearest first (self-edge first with set_diag)
        row = adj[10]
        d = np.sqrt(((xy[row.indices] - xy[10]) ** 2).sum(1))
        assert np.all(np.diff(d) >= 0)


def test_spatial_graph_helpers():
    from metalsinglecell import gr
    A = gr._edges_to_csr(4, np.array([2, 0, 2, 1]), np.array([1, 3, 0, 2]),
                         np.ones(4, np.float32), self_value=np.float32(1.0))
    assert A.indptr.tolist() == [0, 2, 4, 7, 8]
    assert A.indices.tolist() == [0, 3, 1, 2, 2, 1, 0, 3]    # self first, then list order
    S = gr._spectral(A)
    deg = np.asarray(A.sum(axis=0)).ravel()
    np.testing.assert_allclose(S.toarray(), A.toarray() / np.sqrt(np.outer(deg, deg)))
    with pytest.raises(NotImplementedError, match="transform"):
        gr._require_transform("banana")


def test_api_errors_before_any_gpu_work():
    from metalsinglecell import gr, pp, tl
    from metalsinglecell._anndata import AnnData
    a = AnnData(np.zeros((20, 3), np.float32))
    with pytest.raises(NotImplementedError, match="metric"):
        pp.neighbors(a, metric="cosine")
    with pytest.raises(ValueError, match="neighbors"):
        tl.leiden(a, backend="gpu")
    a.obsp["connectivities"] = sp.identity(20, format="csr")
    with pytest.raises(ValueError, match="backend"):
        tl.leiden(a, backend="banana")
    with pytest.raises(ValueError, match="backend"):
        tl.louvain(a, backend="banana")
    with pytest.raises(NotImplementedError, match="variant"):
        tl.louvain(a, backend="gpu", variant="colored")
    with pytest.raises(ValueError, match="commit_prob"):
        tl.louvain(a, backend="gpu", commit_prob=0.0)
    with pytest.raises(ValueError, match="mode"):
        gr.spatial_autocorr(a, mode="Moran")
    with pytest.raises(KeyError, match="spatial"):
        gr.spatial_neighbors_knn(a)
    a.obsm["spatial"] = np.zeros((20, 2))
    with pytest.raises(NotImplementedError, match="transform"):
        gr.spatial_neighbors_knn(a, transform="banana")
    a.obs["lib"] = ["x"] * 20
    with pytest.raises(TypeError, match="categorical"):
        gr.spatial_neighbors_knn(a, library_key="lib")
    with pytest.raises(KeyError, match="nope"):
        gr.spatial_neighbors_knn(a, library_key="nope")
    wide = AnnData(np.zeros((20, 60), np.float32))
    with pytest.raises(NotImplementedError, match="X_pca"):
        pp.neighbors(wide)


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    """Cache keys bucket the largest sample and the total by powers of two; entries persist."""
    from metalsinglecell import autotune
    monkeypatch.setenv("MSC_AUTOTUNE_CACHE", str(tmp_path / "sub" / "t.json"))
    off = np.array([0, 500, 1700, 60_000])
    assert autotune.key_for(off, 30, 15, "bf16") == "knn_batched|bf16|d30|k15|max65536|tot65536"
    assert autotune.lookup(off, 30, 15, "bf16") is None
    autotune.store(autotune.key_for(off, 30, 15, "bf16"),
                   {"splits": 4, "list_len": 8, "ms": {"1x16": 2.0, "4x8": 1.0}})
    assert autotune.lookup(off, 30, 15, "bf16") == (4, 8)
    assert autotune.lookup(off, 30, 15, "fp32") is None
    off2 = np.array([0, 500, 1700, 58_000])          # same buckets -> same entry
    assert autotune.lookup(off2, 30, 15, "bf16") == (4, 8)
    autotune.store(autotune.key_for(off, 30, 40, "bf16"), {"splits": 2, "ms": {"2": 1.0}})
    assert autotune.lookup(off, 30, 40, "bf16") == (2, 0)      # entries without a list length
    # the short list is a candidate only where it holds k (4 lanes x 8 keys) and the splits
    # admit at least 4k candidates per query
    assert autotune.candidates_for(32) == [(s, 16) for s in (1, 2, 4, 8)] + [(4, 8), (8, 8)]
    assert autotune.candidates_for(15) == [(s, 16) for s in (1, 2, 4, 8)] + [(2, 8), (4, 8), (8, 8)]
    assert autotune.candidates_for(33) == [(s, 16) for s in (1, 2, 4, 8)]


def test_choose_representation_follows_the_reference_rule():
    """reference pp.py:732-739: narrow objects use .X even when X_pca exists; wide objects use
    X_pca only when it has n_pcs (or 50) columns, else the reference computes a PCA (raises
    here); n_pcs=0 means .X; an unknown use_rep raises ValueError."""
    from metalsinglecell._anndata import AnnData
    from metalsinglecell.pp import _choose_representation
    rng = np.random.default_rng(0)
    narrow = AnnData(rng.normal(size=(30, 40)).astype(np.float32))
    narrow.obsm["X_pca"] = rng.normal(size=(30, 10)).astype(np.float32)
    rep, key = _choose_representation(narrow, None, None)
    assert key is None and rep.shape == (30, 40) and np.array_equal(rep, narrow.X)
    rep, key = _choose_representation(narrow, "X_pca", None)
    assert key == "X_pca" and rep.shape == (30, 10)
    wide = AnnData(rng.normal(size=(30, 60)).astype(np.float32))
    wide.obsm["X_pca"] = rng.normal(size=(30, 30)).astype(np.float32)
    with pytest.raises(NotImplementedError, match="30 < 50"):
        _choose_representation(wide, None, None)
    rep, key = _choose_representation(wide, None, 20)
    assert key == "X_pca" and rep.shape == (30, 30)
    with pytest.raises(NotImplementedError, match="30 < 40"):
        _choose_representation(wide, None, 40)
    rep, key = _choose_representation(wide, None, 0)
    assert key is None and rep.shape == (30, 60)
    rep, key = _choose_representation(wide, "X", None)
    assert key == "X" and rep.shape == (30, 60)
    with pytest.raises(ValueError, match="Did not find"):
        _choose_representation(wide, "X_umap", None)


def test_autocorr_value_sources():
    """squidpy's extraction modes: var genes (HVG default, explicit names, layers, .raw with
    names missing from .raw dropped), numeric obs columns, obsm columns by position."""
    import types
    import pandas as pd
    from metalsinglecell import gr
    from metalsinglecell._anndata import AnnData
    rng = np.random.default_rng(1)
    X = sp.csr_matrix(rng.poisson(1.0, (12, 5)).astype(np.float32))
    a = AnnData(X)
    a.var_names = ["a", "b", "c", "d", "e"]
    a.var["highly_variable"] = [True, False, True, True, False]
    vals, names = gr._values_from_var(a, None, None, False)
    assert names == ["a", "c", "d"] and np.array_equal(vals, X.toarray()[:, [0, 2, 3]])
    vals, names = gr._values_from_var(a, "e", None, False)
    assert names == ["e"] and vals.shape == (12, 1)
    a.layers["lay"] = X * 2
    vals, _ = gr._values_from_var(a, ["b", "a"], "lay", False)
    assert np.array_equal(vals, 2 * X.toarray()[:, [1, 0]])
    with pytest.raises(AttributeError, match="raw"):
        gr._values_from_var(a, None, None, True)
    raw = types.SimpleNamespace(X=np.arange(36, dtype=np.float32).reshape(12, 3),
                                var_names=pd.Index(["d", "a", "z"]))
    a.raw = raw
    vals, names = gr._values_from_var(a, None, None, True)
    assert names == ["a", "d"] and np.array_equal(vals, raw.X[:, [1, 0]])
    a.obs["num"] = np.arange(12.0)
    a.obs["txt"] = ["x"] * 12
    a.obs["num2"] = np.ones(12)
    vals, names = gr._values_from_obs(a, None, None, False)
    assert names == ["num", "num2"] and vals.shape == (12, 2)
    vals, names = gr._values_from_obs(a, "num2", None, False)
    assert names == ["num2"]
    a.obsm["emb"] = rng.normal(size=(12, 4))
    vals, names = gr._values_from_obsm(a, [3, 1], "emb", False)
    assert names == [3, 1] and np.allclose(vals, a.obsm["emb"][:, [3, 1]].astype(np.float32))
    with pytest.raises(KeyError, match="obsm"):
        gr._values_from_obsm(a, None, "nope", False)
    with pytest.raises(NotImplementedError, match="layers"):
        gr.spatial_autocorr(a, attr="layers")


def test_autocorr_frame_sorting_with_ties_and_nan():
    from metalsinglecell import gr
    W = sp.random(40, 40, density=0.1, random_state=2, format="csr")
    I = np.array([0.2, np.nan, 0.5, 0.2, -0.1])
    df = gr.moran_frame(I, None, W, ["a", "b", "c", "d", "e"])
    assert df.index.tolist() == ["c", "a", "d", "e", "b"]
    df = gr.moran_frame(I, None, W, ["a", "b", "c", "d", "e"], mode="geary")
    assert df.index.tolist() == ["e", "a", "d", "c", "b"]
