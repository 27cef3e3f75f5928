"""Exact kNN with partition-pruned candidate ranges against the CPU oracle.

The pruned layout reorders the rows into partitions, gives each block of 128 queries only the
candidate ranges the triangle inequality cannot rule out, and maps the results back to the
caller's row ids.  Every case here must equal ``knn_ref.c`` bit for bit (indices and float32
distances), like the unpruned path, in both precisions.  The cases cover what the layout
adds: well-separated clusters (most partitions skipped), tiny clusters and far outliers
(singleton partitions, the pooled group, rows whose own partition holds fewer than k points),
no cluster structure (nothing skipped), exact duplicates and equidistant points (ties broken by
the caller's row id, across partitions), 2-D/3-D and 64-D rows, k from 1 to 64, and the smallest
allowed n.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(name):
    from metalsinglecell import synthetic
    rng = np.random.default_rng(11)
    if name == "iso3000":
        X, _ = synthetic.gaussian_mixture_isotropic(n=3000, d=50, n_clusters=8, seed=0)
        return X, 15
    if name == "zipf40k":
        X, _ = synthetic.gaussian_mixture_zipf(n=40_000, d=50, n_clusters=200, seed=2)
        return X, 30
    if name == "aniso20k":
        X, _ = synthetic.gaussian_mixture_anisotropic(n=20_000, d=50, n_clusters=30, seed=1)
        return X, 15
    if name == "uniform":
        return rng.standard_normal((8000, 16)).astype(np.float32), 10
    if name == "dups":  # 1500 points, each three times, shuffled: ties at distance 0 and beyond
        base = rng.normal(0.0, 3.0, (1500, 8)).astype(np.float32)
        X = np.repeat(base, 3, axis=0)
        return np.ascontiguousarray(X[rng.permutation(X.shape[0])]), 10
    if name == "lattice":  # integer lattice: many equal distances, exact in float32
        g = np.stack(np.meshgrid(np.arange(40), np.arange(40), np.arange(4), indexing="ij"), -1)
        X = g.reshape(-1, 3).astype(np.float32)
        return np.ascontiguousarray(X[rng.permutation(X.shape[0])]), 12
    if name == "tissue2d":
        return synthetic.tissue_coordinates(20_000, seed=4), 6
    if name == "d64k64":
        c = rng.normal(0.0, 6.0, (12, 64))
        X = (c[rng.integers(0, 12, 5000)] + rng.standard_normal((5000, 64))).astype(np.float32)
        return X, 64
    if name == "k1":
        X, _ = synthetic.gaussian_mixture_isotropic(n=3000, d=20, n_clusters=8, seed=5)
        return X, 1
    if name == "n256":
        return rng.standard_normal((256, 5)).astype(np.float32), 5
    if name == "outliers":  # clusters plus 300 scattered far points (peeled into singletons)
        c = rng.normal(0.0, 5.0, (10, 30))
        X = c[rng.integers(0, 10, 12_000)] + rng.standard_normal((12_000, 30))
        far = rng.normal(0.0, 60.0, (300, 30))
        X = np.vstack([X, far]).astype(np.float32)
        return np.ascontiguousarray(X[rng.permutation(X.shape[0])]), 15
    raise KeyError(name)


CASES = ["iso3000", "zipf40k", "aniso20k", "uniform", "dups", "lattice", "tissue2d", "d64k64", "k1",
         "n256", "outliers"]


def _oracle(X, k):
    from oracle import knn
    i, _, d = knn.exact_knn_c(X, k)
    return i, d


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("name", CASES)
def test_pruned_knn_bit_exact(name, precision):
    from metalsinglecell import _hip
    X, k = _case(name)
    print(f"[pruned] {name} n={X.shape[0]} d={X.shape[1]} k={k} {precision}", flush=True)
    idx, dist, info = _hip.knn_exact(X, k, precision=precision, return_info=True, prune=True)
    ri, rd = _oracle(X, k)
    print(f"  partitions {info['n_partitions']} groups {info['n_groups']} ranges {info['n_ranges']} "
          f"scanned {info['scanned']:.4f} second pass {info['n_second_pass']} fallback {info['n_fallback']}")
    assert info["n_partitions"] > 0
    assert np.array_equal(idx.cpu().numpy(), ri)
    assert np.array_equal(dist.cpu().numpy(), rd)


def test_pruned_skips_work_on_separated_clusters():
    """The scanned fraction follows the cluster structure: about sum_c p_c^2 of n^2 on
    well-separated clusters, and all of n^2 on data without clusters."""
    from metalsinglecell import _hip
    X, k = _case("iso3000")
    _, _, info = _hip.knn_exact(X, k, precision="bf16", return_info=True, prune=True)
    assert 1 < info["n_groups"] and info["scanned"] < 0.3
    X, k = _case("uniform")
    _, _, info = _hip.knn_exact(X, k, precision="bf16", return_info=True, prune=True)
    assert info["scanned"] > 0.9


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_pruned_equals_unpruned_200k(precision):
    """200,000 rows of the Zipf mixture: the pruned and the full scan return identical arrays,
    and 1,000 random rows equal the oracle."""
    from metalsinglecell import _hip, synthetic
    from oracle import knn
    X, _ = synthetic.gaussian_mixture_zipf(n=200_000, d=50, n_clusters=200, seed=2)
    a_i, a_d, info = _hip.knn_exact(X, 30, precision=precision, return_info=True, prune=True)
    b_i, b_d = _hip.knn_exact(X, 30, precision=precision, prune=False)
    print(f"[pruned 200k] scanned {info['scanned']:.4f} partitions {info['n_partitions']} "
          f"groups {info['n_groups']}", flush=True)
    assert info["scanned"] < 0.3  # 200 separated clusters: about sum_c p_c^2 = 0.19
    assert np.array_equal(a_i.cpu().numpy(), b_i.cpu().numpy())
    assert np.array_equal(a_d.cpu().numpy(), b_d.cpu().numpy())
    rows = np.random.default_rng(0).choice(X.shape[0], 1000, replace=False)
    ri, _, rd = knn.exact_knn_c(X, 30, rows=rows)
    assert np.array_equal(a_i.cpu().numpy()[rows], ri)
    assert np.array_equal(a_d.cpu().numpy()[rows], rd)


def test_prune_on_rejects_shapes_outside_the_layout():
    from metalsinglecell import _hip
    X = np.random.default_rng(0).standard_normal((1000, 70)).astype(np.float32)
    with pytest.raises(RuntimeError, match="pruned"):
        _hip.knn_exact(X, 5, precision="bf16", prune=True)
    X = np.random.default_rng(0).standard_normal((200, 4)).astype(np.float32)
    with pytest.raises(RuntimeError, match="pruned"):
        _hip.knn_exact(X, 5, precision="bf16", prune=True)
