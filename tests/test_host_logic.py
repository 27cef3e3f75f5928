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


def test_spatial_graph_tail():
    from metalsinglecell import gr
    n = 6
    cols = np.array([1, 2, 0, 2, 0, 1, 4, 5, 3, 5, 3, 4])
    indptr = np.arange(n + 1) * 2
    A = sp.csr_matrix((np.ones(12, np.float32), cols, indptr), shape=(n, n))
    D = sp.csr_matrix((np.arange(1, 13, dtype=np.float64), cols, indptr), shape=(n, n))
    a, d = gr._finalize(A, D, None, True, None)
    assert np.all(a.diagonal() == 1.0) and d.diagonal().sum() == 0
    a, d = gr._finalize(A, D, None, False, 50)
    assert d.data.max() <= np.percentile(np.arange(1, 13), 50)
    a, _ = gr._finalize(A, D, "spectral", False, None)
    deg = np.asarray(A.sum(axis=0)).ravel()
    np.testing.assert_allclose(a.toarray(), A.toarray() / np.sqrt(np.outer(deg, deg)))
    with pytest.raises(NotImplementedError, match="transform"):
        gr._transform_adj(A, "banana")


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
    autotune.store(autotune.key_for(off, 30, 15, "bf16"), {"splits": 4, "ms": {"1": 2.0, "4": 1.0}})
    assert autotune.lookup(off, 30, 15, "bf16") == 4
    assert autotune.lookup(off, 30, 15, "fp32") is None
    off2 = np.array([0, 500, 1700, 58_000])          # same buckets -> same entry
    assert autotune.lookup(off2, 30, 15, "bf16") == 4


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
