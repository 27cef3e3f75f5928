"""Serving loop: results identical alone and in the mix; HBM budget admission; failures isolated."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(sizes, G=60, seed=0):
    from metalsinglecell import synthetic
    from metalsinglecell.serve import SpatialJob
    out = []
    for i, n in enumerate(sizes):
        xy, X = synthetic.spatial_job(int(n), n_genes=G, seed=seed + i)
        out.append(SpatialJob(coords=xy, counts=X, n_perms=50, seed=i, job_id=i))
    return out


def _same(a, b):
    return (np.array_equal(a["idx"], b["idx"]) and np.array_equal(a["I"], b["I"], equal_nan=True)
            and np.array_equal(a["sims"], b["sims"], equal_nan=True) and a["moments"] == b["moments"])


def test_alone_equals_in_the_mix():
    import torch
    from metalsinglecell.serve import Server, run_spatial_job
    jobs = _jobs([3000, 20000, 1500, 60000, 8000, 400, 30000, 5000])
    alone = [run_spatial_job(j) for j in jobs]
    torch.cuda.synchronize()
    srv = Server(max_slots=4)
    futs = [srv.submit(j) for j in jobs]
    res = [f.wait(120) for f in futs]
    srv.close()
    for a, b in zip(alone, res):
        assert _same(a, b)
    st = srv.stats()
    assert st["jobs"] == len(jobs) and st["p99_s"] is not None


def test_result_matches_oracle():
    import scipy.sparse as sp
    from metalsinglecell.serve import run_spatial_job
    from oracle import grid, moran
    job = _jobs([2500], G=20)[0]
    r = run_spatial_job(job, permutations="numpy")
    A, _ = grid.knn_graph(job.coords, 6)
    assert np.array_equal(r["idx"].ravel(), A.indices)
    perms = moran.permutations(job.n, job.n_perms, job.seed)
    I, sims = moran.moran_dense(job.counts.toarray(), moran.row_normalize(A), perms)
    np.testing.assert_allclose(r["I"], I, atol=1e-6, rtol=0)
    np.testing.assert_allclose(r["sims"], sims, atol=1e-6, rtol=0)
    s = moran.graph_moments(moran.row_normalize(A))
    np.testing.assert_allclose(r["moments"], s, rtol=1e-12)


def test_budget_rejects_oversized_jobs_and_failures_stay_isolated():
    from metalsinglecell.serve import Server
    jobs = _jobs([2000, 50000, 3000])
    small = jobs[0].footprint()
    srv = Server(hbm_budget=max(jobs[0].footprint(), jobs[2].footprint()) + 1, max_slots=2)
    bad = _jobs([1000])[0]
    bad.coords = bad.coords.copy()
    bad.coords[0, 0] = np.nan                 # rejected by the kernel's host-side check
    futs = [srv.submit(j) for j in jobs] + [srv.submit(bad)]
    with pytest.raises(MemoryError):
        futs[1].wait(60)
    with pytest.raises(RuntimeError, match="finite"):
        futs[3].wait(60)
    r0, r2 = futs[0].wait(60), futs[2].wait(60)
    srv.close()
    assert r0["I"].shape == (60,) and r2["I"].shape == (60,)
    assert small > 0
