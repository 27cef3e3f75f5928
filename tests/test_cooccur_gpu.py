"""HIP co-occurrence pair histogram (msc_co_occurrence) against the oracle, and the
gr.co_occurrence contract of the reference (tests/test_gr_contracts.py:32-66)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _labelled(n=300, k=4, seed=0):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0, 1000, (n, 2))
    lab = rng.integers(0, k, n).astype(str)
    return xy, lab


def _check(xy, codes, K, thr2):
    from metalsinglecell import _hip
    from oracle import cooccur
    h = _hip.co_occurrence_hist(xy.astype(np.float32), codes, K, thr2).cpu().numpy()
    ref = cooccur.pair_hist(xy, codes, K, thr2)
    assert np.array_equal(h, ref)
    return h


def test_hist_reference_fixture():
    from oracle import cooccur
    xy, lab = _labelled()
    cats = np.unique(lab)
    thr = cooccur.thresholds(xy.astype(np.float32).astype(np.float64), 20)
    _check(xy, np.searchsorted(cats, lab), len(cats), (thr[1:] ** 2).astype(np.float32))


@pytest.mark.parametrize("n,K,L,dim", [(5003, 12, 49, 2), (2600, 30, 9, 3), (700, 1, 1, 2)])
def test_hist_sizes(n, K, L, dim):
    rng = np.random.default_rng(n)
    xy = rng.uniform(0, 500, (n, dim))
    codes = rng.integers(0, K, n)
    codes[:K] = np.arange(K)
    thr2 = (np.linspace(0, 250, L + 1)[1:] ** 2).astype(np.float32)
    _check(xy, codes, K, thr2)


def test_hist_lattice_ties_and_duplicates():
    """Thresholds exactly on lattice distances (d2 == thr2 falls in that bin) and coincident
    points (d2 = 0, first bin); uneven code sizes including a single-cell code."""
    r, c = np.divmod(np.arange(900), 30)
    xy = np.c_[c * 10.0, r * 10.0]
    xy[450:460] = xy[0:10]
    codes = (r * 7 + c) % 5
    codes[899] = 5
    thr2 = np.array([0.0, 100.0, 200.0, 400.0, 2500.0], np.float32)
    h = _check(xy, codes, 6, thr2)
    assert h[:, :, 0].sum() == 20
    # evenly spaced distances (the arithmetic bin estimate) landing exactly on lattice distances,
    # with and without a zero first threshold
    for thr in (np.linspace(0.0, 50.0, 6)[1:], np.linspace(0.0, 40.0, 5), np.linspace(5.0, 95.0, 19)):
        _check(xy, codes, 6, (thr ** 2).astype(np.float32))


def test_gr_co_occurrence_contract():
    import pandas as pd
    from metalsinglecell import gr
    from metalsinglecell._anndata import AnnData
    from oracle import cooccur
    xy, lab = _labelled()
    a = AnnData(np.zeros((300, 2), np.float32))
    a.obsm["spatial"] = xy
    a.obs["ct"] = pd.Categorical(lab)
    gr.co_occurrence(a, "ct", interval=20)
    res = a.uns["ct_co_occurrence"]
    assert len(res["interval"]) == 20 and res["occ"].shape == (4, 4, 19)
    ref = cooccur.co_occurrence(xy, lab, n_intervals=20)
    np.testing.assert_allclose(res["interval"], ref["interval"], rtol=0, atol=0)
    np.testing.assert_allclose(res["occ"], ref["occ"], rtol=1e-12, atol=0)
    thr = np.linspace(50.0, 400.0, 8)
    gr.co_occurrence(a, "ct", interval=thr)
    assert np.allclose(a.uns["ct_co_occurrence"]["interval"], thr)
    assert a.uns["ct_co_occurrence"]["occ"].shape[-1] == 7
    a.obs["plain"] = a.obs["ct"].astype(str)
    with pytest.raises(TypeError, match="categorical"):
        gr.co_occurrence(a, "plain")
    del a.uns["ct_co_occurrence"]
    out = gr.co_occurrence(a, "ct", interval=10, copy=True)
    assert isinstance(out, tuple) and len(out) == 2 and "ct_co_occurrence" not in a.uns
