"""Leiden of many samples in one block-diagonal run (msc_leiden_batched) against msc_leiden on
each sample alone and against the oracle: identical labels and modularity per sample, for
uneven sample sizes, several seeds / resolutions / iteration counts, and samples that stop
at different levels."""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _samples(sizes, seed=3, d=20, k=15):
    from metalsinglecell import _hip
    from metalsinglecell.batch import BatchGraph
    rng = np.random.default_rng(seed)
    xs = []
    for m in sizes:
        c = int(rng.integers(2, 12))
        centers = rng.normal(0.0, 5.0, (c, d))
        xs.append((centers[rng.integers(0, c, m)] + rng.normal(0.0, 1.0, (m, d))).astype(np.float32))
    bg = BatchGraph([x.shape[0] for x in xs], d, k=k, precision="bf16", splits=1, capture=False)
    bg.run(np.concatenate(xs))
    return bg


def _block(bg, s):
    a, b = int(bg.offsets[s]), int(bg.offsets[s + 1])
    ip = bg.indptr.cpu().numpy()
    e0, e1 = ip[a], ip[b]
    return (ip[a:b + 1] - e0, bg.indices[e0:e1].cpu().numpy() - a, bg.data[e0:e1].cpu().numpy(), b - a)


@pytest.mark.parametrize("res,seed,iters", [(1.0, 0, 2), (0.5, 7, 1), (2.0, 3, 3)])
def test_batched_equals_single_runs(res, seed, iters):
    from metalsinglecell import _hip
    bg = _samples([700, 64, 2500, 300, 1200, 90, 4000, 40])
    lab, q = _hip.leiden_batched(bg.indptr, bg.indices, bg.data, bg.offsets, resolution=res,
                                 random_state=seed, n_iterations=iters)
    lab = lab.cpu().numpy()
    for s in range(bg.sizes.size):
        ip, ind, dat, m = _block(bg, s)
        ref, rq, _ = _hip.leiden(ip, ind, dat, m, resolution=res, random_state=seed, n_iterations=iters)
        a, b = int(bg.offsets[s]), int(bg.offsets[s + 1])
        assert np.array_equal(lab[a:b], ref.cpu().numpy()), f"sample {s}"
        assert q[s] == rq, f"sample {s}: {q[s]} vs {rq}"


def test_batched_matches_oracle():
    from metalsinglecell import _hip
    from oracle import leiden
    bg = _samples([1500, 200, 900], seed=5)
    lab, q = _hip.leiden_batched(bg.indptr, bg.indices, bg.data, bg.offsets)
    lab = lab.cpu().numpy()
    for s in range(3):
        ip, ind, dat, m = _block(bg, s)
        R = sp.csr_matrix((dat, ind, ip), shape=(m, m))
        ref = leiden.leiden(R, resolution=1.0, random_state=0, n_iterations=2)
        a, b = int(bg.offsets[s]), int(bg.offsets[s + 1])
        assert np.array_equal(lab[a:b], ref)
        assert q[s] == pytest.approx(leiden.modularity_of_labels(R, ref), rel=1e-12)


def test_batched_single_sample_equals_msc_leiden():
    from metalsinglecell import _hip
    bg = _samples([3000], seed=9)
    lab, q = _hip.leiden_batched(bg.indptr, bg.indices, bg.data, bg.offsets)
    nnz = int(bg.indptr[-1].item())
    ref, rq, _ = _hip.leiden(bg.indptr, bg.indices[:nnz], bg.data[:nnz], 3000)
    assert np.array_equal(lab.cpu().numpy(), ref.cpu().numpy()) and q[0] == rq
