"""Batched form: one launch over uneven samples equals per-sample calls, for every split
count, on every HIP-graph replay, through Leiden."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _samples(seed=3, sizes=(200, 1000, 128, 3000, 777, 64), d=30):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        c = int(rng.integers(3, 12))
        centers = rng.normal(0.0, 5.0, (c, d))
        out.append((centers[rng.integers(0, c, n)] + rng.normal(0.0, 1.0, (n, d))).astype(np.float32))
    return out


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("splits,list_len", [(1, 0), (4, 0), (1, 8), (4, 8)])
def test_batched_knn_equals_per_sample(precision, splits, list_len):
    from metalsinglecell import _hip
    S = _samples()
    X = np.concatenate(S)
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in S])])
    idx, dist, info = _hip.knn_batched(X, off, 15, precision=precision, splits=splits, list_len=list_len,
                                       return_info=True)
    idx, dist = idx.cpu().numpy(), dist.cpu().numpy()
    assert info["n_splits"] == splits and info["list_len"] == 4 * (list_len or 16)
    for s, x in enumerate(S):
        ri, rd = _hip.knn_exact(x, 15, precision=precision)
        a, b = off[s], off[s + 1]
        assert np.array_equal(idx[a:b] - a, ri.cpu().numpy())
        assert np.array_equal(dist[a:b], rd.cpu().numpy())


def test_batched_knn_matches_oracle():
    from metalsinglecell import _hip
    from oracle import knn
    S = _samples(seed=4, sizes=(700, 150, 1300))
    X = np.concatenate(S)
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in S])])
    idx, dist = _hip.knn_batched(X, off, 12, precision="bf16", splits=2)
    idx, dist = idx.cpu().numpy(), dist.cpu().numpy()
    for s, x in enumerate(S):
        ri, _, rd = knn.exact_knn_c(x, 12)
        a, b = off[s], off[s + 1]
        assert np.array_equal(idx[a:b] - a, ri) and np.array_equal(dist[a:b], rd)


def test_graph_stage_replays_and_blocks_equal_single_sample():
    """The captured kNN + connectivities stage: each replay (with eager work and syncs in
    between, and new input data) equals the eager single-sample results block by block."""
    import torch
    from metalsinglecell import _hip
    from metalsinglecell.batch import BatchGraph
    S1, S2 = _samples(seed=5), _samples(seed=6)
    sizes = [s.shape[0] for s in S1]
    bg = BatchGraph(sizes, 30, k=15, precision="bf16", splits=2, capture=True)
    assert bg.graph is not None
    for rep, S in enumerate([S1, S2, S1]):
        idx, dist, (ip, ind, dat) = bg.run(np.concatenate(S))
        torch.cuda.synchronize()
        ip_h, ind_h, dat_h = ip.cpu().numpy(), ind.cpu().numpy(), dat.cpu().numpy()
        _ = torch.ones(1000, device="cuda").sum().item()        # eager work between replays
        for s, x in enumerate(S):
            a, b = bg.offsets[s], bg.offsets[s + 1]
            ri, rd = _hip.knn_exact(x, 15, precision="bf16")
            assert np.array_equal(idx[a:b].cpu().numpy() - a, ri.cpu().numpy()), (rep, s)
            (rip, rind, rdat), _ = _hip.fuzzy_connectivities(ri, rd)
            e0, e1 = ip_h[a], ip_h[b]
            assert np.array_equal(ip_h[a:b + 1] - e0, rip.cpu().numpy()), (rep, s)
            assert np.array_equal(ind_h[e0:e1] - a, rind.cpu().numpy()), (rep, s)
            assert np.array_equal(dat_h[e0:e1], rdat.cpu().numpy()), (rep, s)


def test_batched_leiden_equals_per_sample_and_oracle():
    import scipy.sparse as sp
    from metalsinglecell import _hip
    from metalsinglecell.batch import neighbors_leiden_batched
    from oracle import leiden
    S = _samples(seed=7, sizes=(600, 150, 2000, 333))
    labels, qs, bg = neighbors_leiden_batched(S, k=15, precision="bf16", random_state=0, workers=3,
                                              autotune_splits=False, splits=2)
    for s, x in enumerate(S):
        ri, rd = _hip.knn_exact(x, 15, precision="bf16")
        (ip, ind, dat), _ = _hip.fuzzy_connectivities(ri, rd)
        lab, q, _ = _hip.leiden(ip, ind, dat, x.shape[0], random_state=0)
        assert np.array_equal(labels[s].cpu().numpy(), lab.cpu().numpy()) and qs[s] == q
        n = x.shape[0]
        P = sp.csr_matrix((dat.cpu().numpy(), ind.cpu().numpy(), ip.cpu().numpy()), shape=(n, n))
        ref = leiden.leiden(P, resolution=1.0, random_state=0, n_iterations=2)
        assert np.array_equal(lab.cpu().numpy().astype(np.int64), ref)


def test_autotuner_picks_and_caches(tmp_path, monkeypatch):
    from metalsinglecell import autotune
    from metalsinglecell.batch import BatchGraph
    monkeypatch.setenv("MSC_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    S = _samples(seed=8)
    X = np.concatenate(S)
    sizes = [s.shape[0] for s in S]
    bg = BatchGraph(sizes, 30, k=15, precision="bf16", tune_on=X, capture=False)
    assert (bg.splits, bg.list_len) in autotune.candidates_for(15)
    assert autotune.lookup(bg.offsets, 30, 15, "bf16") == (bg.splits, bg.list_len)
    assert len(autotune.load()[autotune.key_for(bg.offsets, 30, 15, "bf16")]["ms"]) == 7
    bg2 = BatchGraph(sizes, 30, k=15, precision="bf16", capture=False)   # from the cache
    assert (bg2.splits, bg2.list_len) == (bg.splits, bg.list_len)
