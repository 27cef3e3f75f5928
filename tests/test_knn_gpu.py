"""HIP exact kNN against the CPU oracle (bit-exact indices, exact float32 distances)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(X, k, precision):
    from metalsinglecell import _hip
    idx, dist, info = _hip.knn_exact(X, k, precision=precision, return_info=True)
    return idx.cpu().numpy().astype(np.int64), dist.cpu().numpy(), info


def _oracle(X, k):
    from oracle import knn
    i, d2, d = knn.exact_knn_c(X, k) if X.shape[0] > 2000 else knn.exact_knn(X, k)
    return i, d


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_knn_mixture_3000x50_k15_bit_exact(c1_data, precision):
    X, _ = c1_data
    idx, dist, info = _run(X, 15, precision)
    ri, rd = _oracle(X, 15)
    print(precision, info)
    assert np.array_equal(idx, ri)
    assert np.array_equal(dist, rd)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("k", [15, 33, 50])
def test_knn_gaussian_500x20_all_k(precision, k):
    """tests/test_postfix_fixes.py:108-121 of the reference asks recall >= 0.99 against
    numpy brute force at k in {15, 33, 50}; the exact path must give recall 1."""
    rng = np.random.default_rng(0)
    X = rng.standard_normal((500, 20)).astype(np.float32)
    idx, dist, _ = _run(X, k, precision)
    ri, rd = _oracle(X, k)
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)
    D = np.sqrt(((X[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1))
    ref = np.argsort(D, axis=1, kind="stable")[:, :k]
    recall = np.mean([len(set(idx[i]) & set(ref[i])) / k for i in range(500)])
    assert recall == 1.0


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("d", [12, 50])
@pytest.mark.parametrize("n", [8, 64, 65, 129, 192, 193, 257])
def test_knn_candidate_tile_counts(precision, d, n):
    """1 to 5 candidate tiles of 64 rows: a single (odd) tile, whole and partial last tiles, the
    pair loop with and without the odd last tile, and tile loads clamped past the end."""
    rng = np.random.default_rng(n + d)
    X = rng.standard_normal((n, d)).astype(np.float32)
    idx, dist, _ = _run(X, 5, precision)
    ri, rd = _oracle(X, 5)
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_knn_duplicates_self_first(precision):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((60, 10)).astype(np.float32)
    X = np.vstack([X, X[:5], X[:5]]).astype(np.float32)
    idx, dist, _ = _run(X, 10, precision)
    ri, rd = _oracle(X, 10)
    assert np.array_equal(idx[:, 0], np.arange(X.shape[0]))
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_knn_large_coordinates_273(precision):
    """Reference KAT tests/test_spatial_neighbors.py:377-386: spots 273 apart stay 273."""
    base = np.array([[9237.0, 15447.0], [9510.0, 15447.0], [9237.0, 15720.0]])
    for off in (0.0, 20_000.0, 200_000.0):
        X = np.ascontiguousarray(base + off, dtype=np.float32)
        _, dist, _ = _run(X, 2, precision)
        assert abs(float(dist[0, 1]) - 273.0) < 1e-3


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_knn_candidate_splits_20000(precision):
    """n = 20000 runs the select kernel with several candidate splits; sampled rows checked."""
    from metalsinglecell import synthetic
    from oracle import knn
    X, _ = synthetic.gaussian_mixture_anisotropic(n=20_000, seed=1)
    idx, dist, info = _run(X, 15, precision)
    assert info["n_splits"] > 1
    rows = np.random.default_rng(0).choice(X.shape[0], 1500, replace=False)
    ri, _, rd = knn.exact_knn_c(X, 15, rows=rows)
    assert np.array_equal(idx[rows], ri) and np.array_equal(dist[rows], rd)


def test_knn_distances_csr_matches_scipy(c1_data):
    import scipy.sparse as sp
    from metalsinglecell import _hip
    from oracle import knn
    X, _ = c1_data
    i, d = _hip.knn_exact(X, 15)
    ip, ind, dat = _hip.knn_distances_csr(i, d)
    n = X.shape[0]
    M = sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))
    R = knn.distances_csr(i.cpu().numpy(), d.cpu().numpy(), n)
    assert np.array_equal(M.indptr, R.indptr) and np.array_equal(M.indices, R.indices)
    assert np.array_equal(M.data, R.data)


def test_knn_certificate_rescue_paths():
    """Widely separated clusters make the query's BF16 rounding residual (~|x| 2^-9) exceed
    the neighbour gaps: the certificate fails in both select passes and those rows go to
    brute force.  Huge offsets with tiny spreads defeat FP32 ranking the same way.  Either
    way the result must equal the oracle's exactly."""
    from oracle import knn
    rng = np.random.default_rng(5)
    centers = rng.normal(0, 2000.0, (20, 50))
    X = (centers[rng.integers(0, 20, 3000)] + rng.normal(0, 1.0, (3000, 50))).astype(np.float32)
    idx, dist, info = _run(X, 15, "bf16")
    ri, _, rd = knn.exact_knn_c(X, 15)
    assert info["n_second_pass"] > 0
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)
    centers = rng.normal(0, 1e6, (10, 8))
    X = (centers[rng.integers(0, 10, 1500)] + rng.normal(0, 1e-2, (1500, 8))).astype(np.float32)
    for prec in ("bf16", "fp32"):
        idx, dist, info = _run(X, 10, prec)
        ri, _, rd = knn.exact_knn_c(X, 10)
        assert info["n_fallback"] > 0
        assert np.array_equal(idx, ri) and np.array_equal(dist, rd)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_knn_lane_partition_rescued_by_wide_pass(precision):
    """In the first select pass a query's candidates are split over lists by index: FP32
    (16x16 tiles) keeps 4 lists of 16, candidate j in list (j mod 16) div 4; BF16 (32x32
    tiles) keeps 2 lists of 32, candidate j in list (j mod 8) div 4.  A query whose k nearest
    neighbours all fall in one list overflows it (FP32: k = 20 with j mod 16 < 4; BF16:
    k = 40 with j mod 8 < 4) and fails the certificate.  The wide pass splits those candidates
    again over 8 lists of 16 by (j mod 16) div 4 and (j mod 64) div 32 and must certify every
    such row without brute force.

    Every point sits in a tight cluster of k + 1 (its k neighbours are its cluster), so no
    near-ties exist across clusters: one group of clusters uses only the indices of one list,
    the other group the remaining indices."""
    from oracle import knn
    rng = np.random.default_rng(11)
    if precision == "fp32":
        n, dim, k, period, width = 2688, 8, 20, 16, 4
    else:
        n, dim, k, period, width = 3280, 8, 40, 8, 4
    j = np.arange(n)
    groups = [j[j % period < width], j[j % period >= width]]
    n_cl = [len(g) // (k + 1) for g in groups]
    assert all(len(g) == m * (k + 1) for g, m in zip(groups, n_cl))
    centers = rng.normal(0.0, 100.0, (sum(n_cl), dim))
    X = np.empty((n, dim))
    c = 0
    for g, m in zip(groups, n_cl):
        for t in range(m):
            X[g[(k + 1) * t: (k + 1) * (t + 1)]] = centers[c] + rng.normal(0.0, 1.0, (k + 1, dim))
            c += 1
    X = X.astype(np.float32)
    idx, dist, info = _run(X, k, precision)
    ri, _, rd = knn.exact_knn_c(X, k)
    assert info["n_second_pass"] >= (k + 1) * n_cl[0]
    assert info["n_fallback"] == 0
    assert np.array_equal(idx, ri) and np.array_equal(dist, rd)


@pytest.mark.parametrize("d", [1, 6, 7, 14, 22, 30, 38, 46, 50, 61, 62, 63, 77, 93, 109, 125])
def test_knn_bf16_every_k_step_shape(d):
    """BF16 operand rows of 2d + 3 values: up to d = 62 in 16-element steps of the 32x32x16
    MFMA (d = 1..62 covers every step count 1..8, several at a step's last full d), beyond in
    32-element steps of the 16x16x32 MFMA (5..8 steps); FP32 rows of d + 1 values in
    8-element steps; both precisions equal the exact oracle."""
    from metalsinglecell import _hip
    rng = np.random.default_rng(d)
    centers = rng.normal(0, 4.0, (12, d))
    X = (centers[rng.integers(0, 12, 1500)] + rng.normal(0, 1.0, (1500, d))).astype(np.float32)
    ri, rd = _oracle(X, 10)
    for precision in ("bf16", "fp32"):
        idx, dist = _hip.knn_exact(X, 10, precision=precision)
        assert np.array_equal(idx.cpu().numpy().astype(np.int64), ri), (d, precision)
        assert np.array_equal(dist.cpu().numpy(), rd), (d, precision)
