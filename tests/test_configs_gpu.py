"""The named workloads at their real sizes, checked against the oracle on the device.

* 100,000 cells x 50 PCs, 30 anisotropic clusters with log-normal sizes (seed 1), k = 15:
  kNN of every row, the connectivities CSR and the Leiden labels bit-identical to the oracle
  (``knn_ref.c``, ``oracle/fuzzy.py``, ``oracle/leiden.py``).
* 1,000,000 cells x 50 PCs, 200 Zipf(1.5) clusters (seed 2), k = 30, BF16 + exact re-rank and
  FP32: 2,000 random rows bit-exact against ``knn_ref.c`` (recall@k = 1 and identical
  distances), the full BF16 result equal to the full FP32 result, the returned distances
  within 1e-5 relative of float64 direct differences, the full connectivities bit-identical
  to the oracle on the same kNN, Leiden's modularity recomputed by the oracle on the device
  CSR, and ARI >= 0.95 against the planted clusters.
* 256 samples of 500-60,000 cells x 30 PCs (seed 3) through the batched form: the smallest,
  a middle and the largest sample each equal to the oracle's kNN, connectivities and Leiden.

The reference checks recall on its real pipeline (tests/test_gpu_parity.py:80-95,
tests/test_postfix_fixes.py:108-121); these are the same checks at the workload sizes.
"""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

_CACHE = {}


def _csr(t, n):
    ip, ind, dat = (x.cpu().numpy() for x in t)
    return sp.csr_matrix((dat, ind, ip), shape=(n, n))


def _same_csr(P, R):
    return (np.array_equal(P.indptr, R.indptr) and np.array_equal(P.indices, R.indices)
            and np.array_equal(P.data, R.data))


# ---------------------------------------------------------------- 100k anisotropic mixture
def _c2():
    if "c2" not in _CACHE:
        from metalsinglecell import synthetic
        from oracle import knn
        X, lab = synthetic.gaussian_mixture_anisotropic(n=100_000, d=50, n_clusters=30, seed=1)
        print(f"[c2] oracle kNN 100000 x 50, k=15", flush=True)
        ri, _, rd = knn.exact_knn_c(X, 15)
        _CACHE["c2"] = (X, lab, ri, rd)
    return _CACHE["c2"]


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_c2_knn_every_row_bit_exact(precision):
    from metalsinglecell import _hip
    X, _, ri, rd = _c2()
    idx, dist, info = _hip.knn_exact(X, 15, precision=precision, return_info=True)
    assert np.array_equal(idx.cpu().numpy(), ri)
    assert np.array_equal(dist.cpu().numpy(), rd)
    print(f"[c2] {precision}: second pass {info['n_second_pass']}, fallback {info['n_fallback']}")


def test_c2_connectivities_and_leiden_bit_identical():
    from metalsinglecell import _hip
    from oracle import fuzzy, leiden
    X, lab, ri, rd = _c2()
    n = X.shape[0]
    idx, dist = _hip.knn_exact(X, 15, precision="bf16")
    conn, _ = _hip.fuzzy_connectivities(idx, dist)
    P = _csr(conn, n)
    R, _, _ = fuzzy.fuzzy_simplicial_set(ri, rd)
    assert _same_csr(P, R)
    labels, q, _ = _hip.leiden(*conn, n, resolution=1.0, random_state=0, n_iterations=2)
    print("[c2] oracle Leiden", flush=True)
    ref = leiden.leiden(R, resolution=1.0, random_state=0, n_iterations=2)
    assert np.array_equal(labels.cpu().numpy(), ref)
    assert q == pytest.approx(leiden.modularity_of_labels(R, ref), rel=1e-12)
    from sklearn.metrics import adjusted_rand_score
    assert adjusted_rand_score(lab, ref) >= 0.95


# ---------------------------------------------------------------- 1M Zipf mixture
def _c3():
    if "c3" not in _CACHE:
        import torch
        from metalsinglecell import _hip, synthetic
        X, lab = synthetic.gaussian_mixture_zipf(n=1_000_000, d=50, n_clusters=200, a=1.5, seed=2)
        Xd = torch.from_numpy(X).to(_hip.device())
        out = {}
        for prec in ("bf16", "fp32"):
            idx, dist, info = _hip.knn_exact(Xd, 30, precision=prec, return_info=True)
            out[prec] = (idx.cpu().numpy(), dist.cpu().numpy(), info)
            del idx, dist
        rows = np.sort(np.random.default_rng(12345).choice(X.shape[0], 2000, replace=False))
        _CACHE["c3"] = (X, lab, out, rows)
    return _CACHE["c3"]


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_c3_knn_sampled_rows_bit_exact(precision):
    from oracle import knn
    X, _, out, rows = _c3()
    idx, dist, info = out[precision]
    ri, _, rd = knn.exact_knn_c(X, 30, rows=rows)
    recall = np.mean([np.intersect1d(a, b).size / 30.0 for a, b in zip(idx[rows], ri)])
    print(f"[c3] {precision}: recall@30 = {recall:.6f} on {rows.size} rows; second pass "
          f"{info['n_second_pass']}, fallback {info['n_fallback']}")
    assert np.array_equal(idx[rows], ri)
    assert np.array_equal(dist[rows], rd)
    assert recall == 1.0


def test_c3_bf16_equals_fp32_everywhere_and_distances_vs_float64():
    X, _, out, rows = _c3()
    bi, bd, _ = out["bf16"]
    fi, fd, _ = out["fp32"]
    assert np.array_equal(bi, fi) and np.array_equal(bd, fd)
    X64 = X.astype(np.float64)
    diff = X64[rows][:, None, :] - X64[bi[rows]]
    d64 = np.sqrt((diff * diff).sum(-1))
    assert np.all(bd[rows][:, 0] == 0.0)
    rel = np.abs(bd[rows][:, 1:].astype(np.float64) - d64[:, 1:]) / d64[:, 1:]
    print(f"[c3] returned distance vs float64: max rel {rel.max():.3e}")
    assert rel.max() <= 1e-5


def test_c3_connectivities_full_and_leiden_quality():
    import torch
    from sklearn.metrics import adjusted_rand_score
    from metalsinglecell import _hip
    from oracle import fuzzy, leiden
    X, lab, out, _ = _c3()
    n = X.shape[0]
    idx, dist, _ = out["bf16"]
    dev = _hip.device()
    conn, _ = _hip.fuzzy_connectivities(torch.from_numpy(idx).to(dev), torch.from_numpy(dist).to(dev))
    P = _csr(conn, n)
    print("[c3] oracle connectivities", flush=True)
    R, _, _ = fuzzy.fuzzy_simplicial_set(idx, dist)
    assert _same_csr(P, R)
    labels, q, _ = _hip.leiden(*conn, n, resolution=1.0, random_state=0, n_iterations=2)
    labels = labels.cpu().numpy()
    q_oracle = leiden.modularity_of_labels(R, labels)
    ari = adjusted_rand_score(lab, labels)
    print(f"[c3] Leiden: {labels.max() + 1} clusters, Q {q:.6f} (oracle {q_oracle:.6f}), ARI {ari:.4f}")
    assert q == pytest.approx(q_oracle, rel=1e-12)
    assert ari >= 0.95


# ---------------------------------------------------------------- 256 batched samples
def test_c4_batched_samples_match_oracle():
    from metalsinglecell import synthetic
    from metalsinglecell.batch import neighbors_leiden_batched
    from oracle import fuzzy, knn, leiden
    samples = synthetic.batched_samples(n_samples=256, lo=500, hi=60_000, d=30, seed=3)
    xs = [x for x, _ in samples]
    sizes = np.array([x.shape[0] for x in xs])
    labels, qs, bg = neighbors_leiden_batched(xs, k=15, precision="bf16", resolution=1.0,
                                              random_state=0, n_iterations=2)
    idx = bg.idx.cpu().numpy()
    dist = bg.dist.cpu().numpy()
    conn = _csr((bg.indptr, bg.indices, bg.data), bg.n)
    order = np.argsort(sizes, kind="stable")
    picks = [int(order[0]), int(order[sizes.size // 2]), int(order[-1])]
    for s in picks:
        a, b = int(bg.offsets[s]), int(bg.offsets[s + 1])
        print(f"[c4] sample {s}: {b - a} cells", flush=True)
        ri, _, rd = knn.exact_knn_c(xs[s], 15)
        assert np.array_equal(idx[a:b] - a, ri) and np.array_equal(dist[a:b], rd)
        R, _, _ = fuzzy.fuzzy_simplicial_set(ri, rd)
        assert _same_csr(conn[a:b, a:b].tocsr(), R)
        ref = leiden.leiden(R, resolution=1.0, random_state=0, n_iterations=2)
        assert np.array_equal(labels[s].cpu().numpy(), ref)
        assert qs[s] == pytest.approx(leiden.modularity_of_labels(R, ref), rel=1e-12)
    assert sizes.min() < 1000 and sizes.max() > 50_000
