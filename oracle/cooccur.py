"""Cluster co-occurrence oracle — TEST INFRASTRUCTURE ONLY.

Restates the reference's ``spatial.co_occurrence`` (src/metalsinglecell/spatial.py:201-258,
histogram kernel :123-146, thresholds :174-187): every ordered pair (i, j), i != j, lands in
the first bin r with ``thr2[r] >= d2`` (thr2 = thresholds[1:]^2 in float32, numpy
``searchsorted`` 'left'), pairs beyond the last threshold are dropped; cumulative counts over
the bins; ``occ[:, c, r] = (counts[c, :, r] / rowsum[c, r]) / P(type)`` with the marginal
``P(type) = rowsum[:, r] / total[r]``.

Deliberate definition shared with the HIP kernel: d2 is the separately rounded float32 sum of
squared coordinate differences (oracle/grid.py), not the Gram form the reference evaluates
(:160-161), which cancels at slide-scale coordinates.
"""

from __future__ import annotations

import numpy as np


def pair_hist(coords, codes, K, thr2, block=512):
    """Per-bin ordered-pair counts (K, K, L) int64 by brute force."""
    X = np.ascontiguousarray(coords, dtype=np.float32)
    n, d = X.shape
    thr2 = np.asarray(thr2, dtype=np.float32)
    L = thr2.size
    codes = np.asarray(codes, dtype=np.int64)
    hist = np.zeros((K, K, L), np.int64)
    for s in range(0, n, block):
        Q = X[s:s + block]
        acc = None
        for t in range(d):
            diff = (X[None, :, t] - Q[:, t, None]).astype(np.float32)
            sq = (diff * diff).astype(np.float32)
            acc = sq if acc is None else (acc + sq).astype(np.float32)
        r = np.searchsorted(thr2, acc, side="left")
        rows = np.arange(Q.shape[0])
        keep = r < L
        keep[rows, s + rows] = False
        qi, j = np.nonzero(keep)
        np.add.at(hist, (codes[s + qi], codes[j], r[qi, j]), 1)
    return hist


def thresholds(spatial, n_intervals):
    coord_sum = spatial.sum(axis=1)
    lo_a, lo_b = np.argpartition(coord_sum, 2)[:2]
    hi = int(np.argmax(coord_sum))
    tmax = float(np.linalg.norm(spatial[lo_a] - spatial[hi])) / 2.0
    tmin = float(np.linalg.norm(spatial[lo_a] - spatial[lo_b]))
    return np.linspace(tmin, tmax, num=n_intervals)


def ratios(hist):
    counts = np.cumsum(hist.astype(np.float64), axis=2)
    K, _, L = counts.shape
    row_sums = counts.sum(axis=0)
    totals = row_sums.sum(axis=0)
    occ = np.zeros((K, K, L))
    for r in range(L):
        probs = row_sums[:, r] / np.maximum(totals[r], 1e-12)
        for c in range(K):
            if row_sums[c, r] != 0.0:
                occ[:, c, r] = (counts[c, :, r] / row_sums[c, r]) / np.maximum(probs, 1e-12)
    return occ


def co_occurrence(coords, labels, n_intervals=50, interval=None):
    X = np.asarray(coords, dtype=np.float32)
    cats = np.unique(np.asarray(labels))
    codes = np.searchsorted(cats, np.asarray(labels))
    thr = (np.sort(np.asarray(interval, np.float64)) if interval is not None
           else thresholds(np.asarray(X, np.float64), n_intervals))
    hist = pair_hist(X, codes, len(cats), (thr[1:] ** 2).astype(np.float32))
    return {"occ": ratios(hist), "interval": thr, "categories": cats, "hist": hist}
