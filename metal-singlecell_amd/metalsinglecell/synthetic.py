"""Seeded synthetic inputs for the neighbor-graph path.

Every generator uses ``numpy.random.default_rng(seed)`` and is fully determined by its
arguments, so a GPU run, the CPU oracle and a benchmark on another host all see the same
cells.  Workload names:

* ``gaussian_mixture_isotropic``   3,000 x 50, 8 clusters, centres N(0, 5^2), sigma 1
* ``gaussian_mixture_anisotropic`` 100,000 x 50, 30 clusters, log-normal sizes, Haar-rotated
                                   per-cluster scales U(0.5, 2)
* ``gaussian_mixture_zipf``        1,000,000 x 50, 200 clusters, Zipf(1.5) sizes, sigma 1
* ``batched_samples``              256 samples, log-uniform sizes in [500, 60000], d = 30
* ``spatial_job``                  2-D tissue coordinates in 10 uniform ellipses plus
                                   negative-binomial counts for Moran's I
"""

from __future__ import annotations

import numpy as np


def _shuffle(rng, X, lab):
    perm = rng.permutation(X.shape[0])
    return np.ascontiguousarray(X[perm]), lab[perm]


def gaussian_mixture_isotropic(n=3000, d=50, n_clusters=8, seed=0):
    """Isotropic mixture: ``centers ~ N(0, 5^2)``, labels uniform, unit within-cluster sigma."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, 5.0, (n_clusters, d))
    lab = rng.integers(0, n_clusters, n)
    X = (centers[lab] + rng.normal(0.0, 1.0, (n, d))).astype(np.float32)
    return X, lab


def _haar_orthogonal(rng, d):
    Z = rng.normal(size=(d, d))
    Q, R = np.linalg.qr(Z)
    return Q * np.sign(np.diag(R))[None, :]


def gaussian_mixture_anisotropic(n=100_000, d=50, n_clusters=30, seed=1):
    """Anisotropic mixture with log-normal cluster weights (mu 0, sigma 1)."""
    rng = np.random.default_rng(seed)
    w = rng.lognormal(0.0, 1.0, n_clusters)
    sizes = rng.multinomial(n, w / w.sum())
    parts, labs = [], []
    for c, m in enumerate(sizes):
        center = rng.normal(0.0, 5.0, d)
        s = rng.uniform(0.5, 2.0, d)
        Q = _haar_orthogonal(rng, d)
        Z = rng.normal(size=(m, d))
        parts.append(center + (Z * s) @ Q.T)
        labs.append(np.full(m, c))
    X = np.concatenate(parts).astype(np.float32)
    return _shuffle(rng, X, np.concatenate(labs))


def gaussian_mixture_zipf(n=1_000_000, d=50, n_clusters=200, a=1.5, seed=2):
    """Isotropic mixture with heavy-tailed (Zipf a=1.5) cluster sizes."""
    rng = np.random.default_rng(seed)
    w = np.arange(1, n_clusters + 1, dtype=np.float64) ** (-a)
    sizes = rng.multinomial(n, w / w.sum())
    centers = rng.normal(0.0, 5.0, (n_clusters, d)).astype(np.float32)
    lab = np.repeat(np.arange(n_clusters), sizes)
    X = centers[lab]
    X += rng.standard_normal((n, d), dtype=np.float32)
    return _shuffle(rng, X, lab)


def batched_samples(n_samples=256, lo=500, hi=60_000, d=30, seed=3):
    """Independent samples with log-uniform sizes, each a 5-40 cluster isotropic mixture."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_samples))).astype(np.int64)
    out = []
    for m in sizes:
        c = int(rng.integers(5, 41))
        centers = rng.normal(0.0, 5.0, (c, d))
        lab = rng.integers(0, c, int(m))
        out.append(((centers[lab] + rng.normal(0.0, 1.0, (int(m), d))).astype(np.float32), lab))
    return out


def tissue_coordinates(n, n_blobs=10, rng=None, seed=4):
    """2-D coordinates with uniform density inside ``n_blobs`` random ellipses (micrometres)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    centers = rng.uniform(0.0, 10_000.0, (n_blobs, 2))
    semi = rng.uniform(500.0, 2_500.0, (n_blobs, 2))
    ang = rng.uniform(0.0, np.pi, n_blobs)
    blob = rng.integers(0, n_blobs, n)
    r = np.sqrt(rng.uniform(0.0, 1.0, n))
    t = rng.uniform(0.0, 2 * np.pi, n)
    u = np.c_[r * np.cos(t) * semi[blob, 0], r * np.sin(t) * semi[blob, 1]]
    c, s = np.cos(ang[blob]), np.sin(ang[blob])
    xy = np.c_[c * u[:, 0] - s * u[:, 1], s * u[:, 0] + c * u[:, 1]] + centers[blob]
    return xy.astype(np.float32)


def nb_counts(n, n_genes=500, density=0.05, theta=2.0, rng=None, seed=4):
    """Negative-binomial counts as CSR (int32 indices, float32 data).

    ``mu_g = LogNormal(0, 1) * s`` with ``s`` chosen by bisection so that the mean nonzero
    fraction under NB(mu, theta), ``1 - (theta / (theta + mu))^theta``, is ``density``.
    """
    import scipy.sparse as sp
    rng = np.random.default_rng(seed) if rng is None else rng
    base = rng.lognormal(0.0, 1.0, n_genes)

    def nz(s):
        mu = base * s
        return float(np.mean(1.0 - (theta / (theta + mu)) ** theta))

    lo, hi = 1e-6, 1e3
    for _ in range(100):
        mid = np.sqrt(lo * hi)
        lo, hi = (mid, hi) if nz(mid) < density else (lo, mid)
    mu = base * np.sqrt(lo * hi)
    q = theta / (theta + mu)                       # NB "success" probability per gene
    p_nz = 1.0 - q ** theta
    rows_all, cols_all, vals_all = [], [], []
    for g in range(n_genes):
        # nonzero positions: geometric gaps (Bernoulli(p_nz) per cell), O(nnz)
        est = int(n * p_nz[g] * 1.2) + 64
        pos = np.cumsum(rng.geometric(p_nz[g], size=est)) - 1
        while pos.size and pos[-1] < n:
            more = np.cumsum(rng.geometric(p_nz[g], size=est)) + pos[-1]
            pos = np.concatenate([pos, more])
        pos = pos[pos < n]
        # zero-truncated NB values by inverse CDF: P(k) = C(k+theta-1, k) q^theta (1-q)^k
        kk = np.arange(1, 400)
        logp = (np.log(q[g]) * theta + kk * np.log1p(-q[g])
                + np.cumsum(np.log((kk + theta - 1) / kk)))
        pk = np.exp(logp) / p_nz[g]
        cdf = np.cumsum(pk)
        cdf[-1] = max(cdf[-1], 1.0)
        vals = 1 + np.searchsorted(cdf, rng.uniform(0.0, cdf[-1], pos.size))
        rows_all.append(pos)
        cols_all.append(np.full(pos.size, g, np.int32))
        vals_all.append(vals.astype(np.float32))
    rows = np.concatenate(rows_all)
    cols = np.concatenate(cols_all)
    vals = np.concatenate(vals_all)
    out = sp.csr_matrix((vals, (rows, cols)), shape=(n, n_genes), dtype=np.float32)
    out.sort_indices()
    out.indices = out.indices.astype(np.int32)
    out.indptr = out.indptr.astype(np.int64)
    return out


def spatial_job(n, n_genes=500, seed=4):
    """One spatial job: coordinates (n, 2) float32 and an (n, n_genes) CSR count matrix."""
    rng = np.random.default_rng(seed)
    xy = tissue_coordinates(n, rng=rng)
    X = nb_counts(n, n_genes, rng=rng)
    return xy, X


def poisson_job_stream(n_jobs, rate=20.0, lo=1_000, hi=2_000_000, seed=4):
    """Arrival times (s) and sizes for a Poisson stream with log-uniform job sizes."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_jobs)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_jobs))).astype(np.int64)
    return np.cumsum(gaps), sizes
