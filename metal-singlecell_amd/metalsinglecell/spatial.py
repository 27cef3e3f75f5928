"""Spatial autocorrelation (squidpy ``gr.spatial_autocorr`` compute), Moran's I and Geary's C
on the MI355X.

Mirrors the reference's ``metalsinglecell.spatial`` (src/metalsinglecell/spatial.py):
``spatial_autocorr(X, connectivity, mode, n_perms, random_state) -> {mode: I, "sims": ...}``
(:357-406) with the statistic and the permutation null computed by csrc/moran.hip (Geary:
:388-393), and the
small host-side statistics over the (genes,) / (perms, genes) results:
``row_normalize`` (:268-279), ``graph_moments`` and ``analytic_pval`` (:282-313, Cliff & Ord
1981 / pysal), ``permutation_pvals`` (:316-331) and ``multiple_testing`` (:334-354).
"""

from __future__ import annotations

import numpy as np

from . import _hip


def row_normalize(g):
    """L1-normalise every row (squidpy ``transformation=True``); zero rows stay zero."""
    import scipy.sparse as sp
    m = sp.csr_matrix(g, dtype=np.float64, copy=True)
    s = np.asarray(np.abs(m).sum(axis=1)).ravel()
    scale = np.zeros_like(s)
    np.divide(1.0, s, out=scale, where=s != 0)
    return sp.csr_matrix(sp.diags(scale) @ m)


def graph_moments(g):
    """pysal's weight moments ``(s0, s1, s2)``: total, half the squared symmetric sum, and
    the summed squared (row + column) totals."""
    import scipy.sparse as sp
    m = sp.csr_matrix(g, dtype=np.float64)
    sym = m + m.T
    s0 = float(m.sum())
    s1 = 0.5 * float(sym.multiply(sym).sum())
    rc = np.asarray(m.sum(axis=1)).ravel() + np.asarray(m.sum(axis=0)).ravel()
    s2 = float(np.dot(rc, rc))
    return s0, s1, s2


def _normal_tail(z):
    """One-sided normal tail toward the sign of z, as ``1 - Phi(z)`` / ``Phi(z)``."""
    from scipy import stats
    return np.where(z > 0, 1.0 - stats.norm.cdf(z), stats.norm.cdf(z))


def analytic_pval(score, g, mode="moran", two_tailed=False):
    """Normality-assumption p-value and variance (Cliff & Ord 1981, as pysal/esda): Moran's
    I around E = -1/(n-1), Geary's C around E = 1 with its own variance."""
    if mode not in ("moran", "geary"):
        raise ValueError(f"mode must be 'moran' or 'geary', got {mode!r}")
    s0, s1, s2 = graph_moments(g)
    n = g.shape[0]
    s02 = s0 * s0
    if mode == "geary":
        var = ((2.0 * s1 + s2) * (n - 1) - 4.0 * s02) / (2.0 * (n + 1) * s02)
        expected = 1.0
    else:
        expected = -1.0 / (n - 1)
        var = (n * n * s1 - n * s2 + 3.0 * s02) / ((n * n - 1.0) * s02) - expected * expected
    with np.errstate(divide="ignore", invalid="ignore"):
        p = _normal_tail((np.asarray(score, np.float64) - expected) / np.sqrt(var))
    return (2.0 * p if two_tailed else p), var


def permutation_pvals(score, sims):
    """Folded two-sided permutation p-value, its z-score p-value and the null variance."""
    P = sims.shape[0]
    ge = (sims >= score[None, :]).sum(axis=0)
    folded = np.minimum(ge, P - ge)
    p_sim = (folded + 1.0) / (P + 1.0)
    mu = sims.mean(axis=0)
    sd = sims.std(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        p_z = _normal_tail((score - mu) / sd)
    return p_sim, p_z, sims.var(axis=0)


def multiple_testing(pvals, method="fdr_bh"):
    """Benjamini-Hochberg step-up (``fdr_bh``) or Bonferroni correction, capped at 1."""
    p = np.asarray(pvals, dtype=np.float64)
    m = p.size
    if method == "bonferroni":
        return np.minimum(p * m, 1.0)
    if method != "fdr_bh":
        raise NotImplementedError(f"corr_method={method!r} is not implemented "
                                  f"(expected 'fdr_bh', 'bonferroni' or None)")
    if m == 0:
        return p.copy()
    order = np.argsort(p, kind="mergesort")
    scaled = p[order] * m / np.arange(1, m + 1)
    adj = np.minimum(np.minimum.accumulate(scaled[::-1])[::-1], 1.0)
    out = np.empty(m)
    out[order] = adj
    return out


def permutations_numpy(n, n_perms, random_state):
    """The reference's permutation stream: ``default_rng(seed).permutation(n)`` in sequence."""
    rng = np.random.default_rng(random_state)
    return np.stack([rng.permutation(n) for _ in range(n_perms)]).astype(np.int32)


def spatial_autocorr(X, connectivity, mode="moran", n_perms=None, random_state=0,
                     permutations="numpy", transformation=False):
    """Per-gene Moran's I or Geary's C and (optionally) its ``(n_perms, genes)`` permutation null.

    ``connectivity`` is used as given (the caller row-normalises, as in the reference),
    unless ``transformation=True``, which normalises on the device.  ``permutations`` picks
    the null's source: ``"numpy"`` draws the reference's ``default_rng(random_state)`` stream
    on the host; ``"device"`` uses a keyed Feistel bijection generated in the kernel.
    """
    import scipy.sparse as sp
    if mode not in ("moran", "geary"):
        raise ValueError(f"mode must be 'moran' or 'geary', got {mode!r}")
    W = sp.csr_matrix(connectivity)
    n = W.shape[0]
    P = int(n_perms or 0)
    perms = None
    if P and permutations == "numpy":
        perms = permutations_numpy(n, P, random_state)
    elif P and permutations != "device":
        raise ValueError("permutations must be 'numpy' or 'device'")
    fn = _hip.moran if mode == "moran" else _hip.geary
    stat, sims, _ = fn(X, W.indptr, W.indices, W.data.astype(np.float64), n_perms=P, perms=perms,
                       seed=random_state, transformation=transformation)
    return {mode: stat.cpu().numpy(), "sims": None if sims is None else sims.cpu().numpy()}


def cooccur_thresholds(spatial):
    """squidpy's ``_find_min_max`` heuristic (spatial.py:174-187 of the reference): anchored on
    the two points with the smallest coordinate sums and half the span to the largest."""
    coord_sum = spatial.sum(axis=1)
    lo_a, lo_b = np.argpartition(coord_sum, 2)[:2]
    hi = int(np.argmax(coord_sum))
    thres_max = float(np.linalg.norm(spatial[lo_a] - spatial[hi])) / 2.0
    thres_min = float(np.linalg.norm(spatial[lo_a] - spatial[lo_b]))
    return thres_min, thres_max


def cooccur_ratios(hist):
    """Cumulative counts and squidpy's conditional normalisation (reference spatial.py:240-258):
    ``occ[:, c, r] = P(type | c, within thr_r) / P(type)``."""
    counts = np.cumsum(np.asarray(hist, dtype=np.float64), axis=2)
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


def co_occurrence(coords, labels, n_intervals=50, interval=None, max_dist=None):
    """Cumulative cluster co-occurrence ratio (squidpy ``gr.co_occurrence``, reference
    spatial.py:201-258): ``{"occ": (K, K, L-1), "interval": thresholds, "categories": cats}``.

    The all-pairs per-bin histogram runs in csrc/cooccur.hip; thresholds, the cumulative sum
    and the normalisation over the small (K, K, L) array run here."""
    X = np.ascontiguousarray(coords, dtype=np.float32)
    labels = np.asarray(labels)
    cats = np.unique(labels)
    code = np.searchsorted(cats, labels).astype(np.int32)
    if interval is not None:
        thr = np.sort(np.asarray(interval, dtype=np.float64))
    else:
        lo, hi = cooccur_thresholds(np.asarray(X, dtype=np.float64))
        if max_dist is not None:
            hi = float(max_dist)
        thr = np.linspace(lo, hi, num=n_intervals)
    if len(thr) <= 1:
        raise ValueError(f"expected interval to be of length >= 2, found {len(thr)}")
    thr2 = (thr[1:] ** 2).astype(np.float32)
    hist = _hip.co_occurrence_hist(X, code, len(cats), thr2).cpu().numpy()
    return {"occ": cooccur_ratios(hist), "interval": thr, "categories": cats}
