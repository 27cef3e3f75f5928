"""Pins the co-occurrence oracle (oracle/cooccur.py): cumulative counts against float64 scipy
distances, pair symmetry, the normalisation on planted and label-shuffled layouts, and the
threshold heuristic's anchors."""

import numpy as np
import pytest
from scipy.spatial.distance import cdist

from oracle import cooccur


def _labelled(n=300, k=4, seed=0):
    """The reference's contract fixture (tests/test_gr_contracts.py:19-29)."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0, 1000, (n, 2))
    lab = rng.integers(0, k, n).astype(str)
    return xy, lab


def test_cumulative_counts_match_float64_distances():
    xy, lab = _labelled()
    res = cooccur.co_occurrence(xy, lab, n_intervals=20)
    thr = res["interval"]
    cats = res["categories"]
    codes = np.searchsorted(cats, lab)
    D = cdist(xy.astype(np.float32).astype(np.float64), xy.astype(np.float32).astype(np.float64))
    np.fill_diagonal(D, np.inf)
    counts = np.cumsum(res["hist"], axis=2)
    for r, t in enumerate(thr[1:]):
        within = D <= t
        for a in range(len(cats)):
            for b in range(len(cats)):
                assert counts[a, b, r] == within[np.ix_(codes == a, codes == b)].sum()


def test_pairs_are_ordered_and_symmetric():
    xy, lab = _labelled(250, 3, 1)
    h = cooccur.co_occurrence(xy, lab, n_intervals=12)["hist"]
    assert np.array_equal(h, h.transpose(1, 0, 2))


def test_planted_clusters_and_shuffled_labels():
    """Three far-apart clumps with one label each: at a radius inside a clump only same-label
    pairs exist, so occ[a, a] = 1 / P(a) and occ[b, a] = 0.  Labels drawn independently of
    position: occ ~ 1 at a wide radius."""
    rng = np.random.default_rng(2)
    sizes = [60, 100, 140]
    xy = np.concatenate([rng.normal(c, 5.0, (m, 2)) for c, m in zip([0, 1000, 2000], sizes)])
    lab = np.repeat(["a", "b", "c"], sizes)
    res = cooccur.co_occurrence(xy, lab, interval=np.array([0.0, 30.0, 60.0]))
    occ = res["occ"]
    for a in range(3):
        r = 0
        # P(type | a within 30) = 1 for a, 0 otherwise; marginal P(a) from the pair counts
        rows = np.cumsum(res["hist"], axis=2).sum(axis=0)[:, r]
        marg = rows / rows.sum()
        assert occ[a, a, r] == pytest.approx(1.0 / marg[a], rel=1e-12)
        assert all(occ[b, a, r] == 0.0 for b in range(3) if b != a)
    xy2 = rng.uniform(0, 1000, (1200, 2))
    lab2 = rng.integers(0, 3, 1200)
    occ2 = cooccur.co_occurrence(xy2, lab2, interval=np.array([0.0, 300.0, 600.0]))["occ"]
    assert np.abs(occ2[:, :, -1] - 1.0).max() < 0.1


def test_threshold_heuristic_anchors():
    """thres_min = distance between the two smallest coordinate sums, thres_max = half the
    distance from the smallest to the largest (squidpy's _find_min_max)."""
    xy = np.array([[0.0, 0.0], [3.0, 4.0], [10.0, 10.0], [1.0, 0.0], [7.0, 1.0]])
    thr = cooccur.thresholds(xy, 5)
    assert thr[0] == pytest.approx(1.0) and thr[-1] == pytest.approx(np.hypot(10, 10) / 2)
    assert len(thr) == 5
