"""Moran's I and Geary's C oracle (float64, dense and sparse forms) — TEST INFRASTRUCTURE ONLY.

Restates src/metalsinglecell/spatial.py:357-406 (statistic and permutation null),
:268-279 (row normalisation), :282-313 (pysal weight moments and the normality p-value,
Cliff & Ord 1981), :316-331 (folded permutation p-values) and :334-354 (Benjamini-Hochberg),
plus the frame assembly of gr.py:455-503.

Statistic: ``Xc = X - mean``, ``denom_g = sum_i Xc_ig^2``, ``Y = W Xc``,
``I_g = (n / S0) * sum_i Xc_ig Y_ig / denom_g``.  Null: for permutation pi drawn from
``np.random.default_rng(seed).permutation(n)`` in sequence, the reference scatters edges from
``pi[src]`` (spatial.py:401-405), which equals ``sim_g = (n/S0) sum_r Xc[pi(r), g] Y[r, g] / denom_g``.
Geary's C (:388-393) is the direct edge sum ``(n-1)/(2 S0) sum_e w_e (Xc_src - Xc_dst)^2 /
denom`` with the same permuted sources.  Everything here is float64.
"""

from __future__ import annotations

import numpy as np


def row_normalize(W):
    import scipy.sparse as sp
    g = sp.csr_matrix(W, copy=True).astype(np.float64)
    s = np.asarray(abs(g).sum(axis=1)).ravel()
    inv = np.divide(1.0, s, out=np.zeros_like(s), where=s != 0)
    return sp.diags(inv) @ g


def permutations(n, n_perms, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.permutation(n) for _ in range(n_perms)]) if n_perms else None


def moran_dense(X, W, perms=None):
    """Dense float64 brute force: ``(I (G,), sims (P, G) or None)``."""
    Wd = W.toarray() if hasattr(W, "toarray") else np.asarray(W)
    Wd = np.asarray(Wd, np.float64)
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    Xc = X - X.mean(axis=0)
    denom = (Xc * Xc).sum(axis=0)
    S0 = Wd.sum()
    Y = Wd @ Xc
    with np.errstate(divide="ignore", invalid="ignore"):
        I = (n / S0) * (Xc * Y).sum(axis=0) / denom
        sims = None
        if perms is not None:
            sims = np.stack([(n / S0) * (Xc[p] * Y).sum(axis=0) / denom for p in perms])
    return I, sims


def moran_scatter(X, W, perms=None):
    """The reference's edge-scatter form (spatial.py:261-265, 372-405) in float64."""
    import scipy.sparse as sp
    coo = sp.coo_matrix(W)
    src, dst, w = coo.row, coo.col, coo.data.astype(np.float64)
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    Xc = X - X.mean(axis=0)
    denom = (Xc * Xc).sum(axis=0)
    S0 = float(w.sum())

    def stat(s):
        out = np.zeros_like(Xc)
        np.add.at(out, s, w[:, None] * Xc[dst])
        with np.errstate(divide="ignore", invalid="ignore"):
            return (n / S0) * (Xc * out).sum(axis=0) / denom

    I = stat(src)
    sims = None if perms is None else np.stack([stat(p[src]) for p in perms])
    return I, sims


def geary_scatter(X, W, perms=None):
    """Geary's C in the reference's edge form (spatial.py:388-393, null :396-406), float64:
    ``C_g = (n-1)/(2 S0) sum_e w_e (Xc[src_e] - Xc[dst_e])^2 / denom_g``; permutation pi
    moves the sources to ``pi[src]``."""
    import scipy.sparse as sp
    coo = sp.coo_matrix(W)
    src, dst, w = coo.row, coo.col, coo.data.astype(np.float64)
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    Xc = X - X.mean(axis=0)
    denom = (Xc * Xc).sum(axis=0)
    S0 = float(w.sum())

    def stat(s):
        de = Xc[s] - Xc[dst]
        with np.errstate(divide="ignore", invalid="ignore"):
            return (n - 1) / (2.0 * S0) * (w[:, None] * de * de).sum(axis=0) / denom

    C = stat(src)
    sims = None if perms is None else np.stack([stat(p[src]) for p in perms])
    return C, sims


def geary_variance(g):
    """Normality variance of Geary's C (Cliff & Ord 1981, as pysal/esda; spatial.py:304-306)."""
    s0, s1, s2 = graph_moments(g)
    n = g.shape[0]
    s02 = s0 * s0
    return ((2 * s1 + s2) * (n - 1) - 4 * s02) / (2 * (n + 1) * s02)


def graph_moments(g):
    s0 = float(g.sum())
    t = g.transpose() + g
    s1 = float(t.multiply(t).sum()) / 2.0
    row = np.asarray(g.sum(axis=1)).ravel()
    col = np.asarray(g.sum(axis=0)).ravel()
    s2 = float(((row + col) ** 2).sum())
    return s0, s1, s2


def analytic_pval(score, g, two_tailed=False):
    from scipy import stats
    s0, s1, s2 = graph_moments(g)
    n = g.shape[0]
    s02 = s0 * s0
    var = (n * n * s1 - n * s2 + 3 * s02) / ((n - 1) * (n + 1) * s02) - (1.0 / (n - 1)) ** 2
    expected = -1.0 / (n - 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        z = (score - expected) / np.sqrt(var)
        p = np.where(z > 0, 1.0 - stats.norm.cdf(z), stats.norm.cdf(z))
    return (p * 2.0 if two_tailed else p), var


def permutation_pvals(score, sims):
    from scipy import stats
    P = sims.shape[0]
    large = (sims >= score).sum(axis=0)
    large = np.where((P - large) < large, P - large, large)
    p_sim = (large + 1.0) / (P + 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        z = (score - sims.mean(axis=0)) / sims.std(axis=0)
        p_z = np.where(z > 0, 1.0 - stats.norm.cdf(z), stats.norm.cdf(z))
    return p_sim, p_z, sims.var(axis=0)


def bh(p):
    """Benjamini-Hochberg step-up by its definition: ``min_{j>=i} p_(j) n / j``, capped at 1."""
    p = np.asarray(p, np.float64)
    n = p.size
    order = np.argsort(p)
    adj_sorted = np.empty(n)
    for i in range(n):
        adj_sorted[i] = min(min(p[order[j]] * n / (j + 1) for j in range(i, n)), 1.0)
    out = np.empty(n)
    out[order] = adj_sorted
    return out
