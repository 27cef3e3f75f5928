"""Neighbor-graph benchmark: cells/sec for exact kNN + connectivities + Leiden on one MI355X.

Workload (default): 1,000,000 cells x 50 PCs float32, 200-cluster Gaussian mixture with
Zipf(1.5) cluster sizes (seed 2), k = 30, Leiden resolution 1.0, seed 0, n_iterations 2.
One step = one full pass from the device-resident embedding to device labels:

    exact kNN (partition-pruned candidate ranges, BF16 MFMA ranking, 64 admitted candidates,
    exact FP32 re-rank)
    -> UMAP fuzzy connectivities -> synchronous parallel Leiden.

``value`` = cells / (mean step time) for that BF16+re-rank form.  The plain FP32 GPU form
(FP32 MFMA ranking) is timed beside it on the same input and reported in ``fp32_form``;
both forms return identical graphs and labels; ``fp32_form.full_scan`` is the FP32 form without
the partition-pruned candidate ranges (every query scans all n rows).  ``roofline`` covers the kernel
with the most time per step, the Leiden move kernel (one launch per pass): algorithmic bytes
8 E_l + 24 n_l per launch over its HIP-event duration, against 8 TB/s HBM.  ``roofline_select``
covers the dominant MFMA kernel (the kNN select kernel): algorithmic flops 2 d per scanned
(query, candidate) pair per launch -- s n^2 pairs, s the scanned fraction of the partition-pruned
layout, 1 without it -- over its HIP-event duration, against the 2.5 PFLOP/s dense BF16 peak.  ``cpu_baseline`` times the project's
CPU restatement (oracle/) of the same three stages on bounded samples on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "metal-singlecell_amd"))
sys.path.insert(0, ROOT)

PEAK = {"bf16": 2500.0, "fp32": 157.3}   # dense MFMA TFLOP/s, MI355X datasheet


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_step(Xd, k, precision, resolution=1.0, seed=0, n_iterations=2, prune=None):
    from metalsinglecell import _hip
    import torch
    t0 = time.perf_counter()
    idx, dist, kinfo = _hip.knn_exact(Xd, k, precision=precision, timing=True, prune=prune)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    (ip, ind, dat), _ = _hip.fuzzy_connectivities(idx, dist)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    labels, q, linfo = _hip.leiden(ip, ind, dat, Xd.shape[0], resolution=resolution,
                                   random_state=seed, n_iterations=n_iterations)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    return labels, {"knn_s": t1 - t0, "conn_s": t2 - t1, "leiden_s": t3 - t2, "q": q,
                    "nnz": int(ind.numel()), "knn": kinfo, "leiden_levels": len(linfo["levels"]),
                    "n_clusters": int(labels.max().item()) + 1}


def timed(Xd, k, precision, steps, warmup, prune=None):
    import torch
    for _ in range(warmup):
        gpu_step(Xd, k, precision, prune=prune)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = []
    labels = None
    for _ in range(steps):
        labels, r = gpu_step(Xd, k, precision, prune=prune)
        recs.append(r)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return total / steps, recs, labels


def move_roofline(Xd, k, select_ms):
    """Roofline of the Leiden move kernel (k_move, one launch per pass), the kernel with the
    most time per step: one more Leiden call on the step's graph with every k_move launch
    bracketed by HIP events on the call's stream.  Algorithmic bytes per launch: 8 E_l + 24 n_l
    (every edge and vertex of the level graph once per pass); HBM traffic per launch from the
    committed FETCH_SIZE / WRITE_SIZE summary of the same kernel and workload."""
    import torch
    from metalsinglecell import _hip
    idx, dist = _hip.knn_exact(Xd, k, precision="bf16")
    (ip, ind, dat), _ = _hip.fuzzy_connectivities(idx, dist)
    n = idx.shape[0]
    _, _, info = _hip.leiden(ip, ind, dat, n, resolution=1.0, random_state=0, n_iterations=2, time_moves=True)
    torch.cuda.synchronize()
    launches = info["move_launches"]
    avg_ms = info["move_ms"] / launches
    bytes_avg = info["move_model_bytes"] / launches
    achieved = bytes_avg / (avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "leiden_move_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    return {"bound": "hbm", "kernel": "k_move (Leiden best move, one launch per pass)", "achieved": achieved,
            "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0, "traffic": traffic,
            "bytes_per_launch": bytes_avg, "avg_launch_ms": avg_ms, "launches_per_step": launches,
            "ms_per_step": info["move_ms"], "select_ms_per_step": select_ms,
            "bytes_model": "8 E_l + 24 n_l per launch (every edge and vertex of the level graph once per pass)"}


def cpu_baseline(X, k, threads, gpu_idx=None):
    """CPU restatement timed on bounded samples of the same workload (extrapolated), with a
    single-thread row; the same sampled kNN rows check the device result (recall@k and
    bit-exact rows against the C oracle)."""
    import numpy as np
    from metalsinglecell import synthetic  # noqa: F401
    from oracle import fuzzy, knn, leiden
    n = X.shape[0]
    rng = np.random.default_rng(0)
    # kNN: queries are independent -> time q queries against all n candidates
    q = 1000 if n >= 200_000 else min(n, 4000)
    rows = np.sort(rng.choice(n, q, replace=False))
    t = time.perf_counter()
    ri, _, rd = knn.exact_knn_c(X, k, rows=rows, threads=threads)
    t_knn = (time.perf_counter() - t) * n / q
    q1 = max(1, q // 10)
    t = time.perf_counter()
    knn.exact_knn_c(X, k, rows=rows[:q1], threads=1)
    t_knn1 = (time.perf_counter() - t) * n / q1
    check = None
    if gpu_idx is not None:
        gi, gd = gpu_idx
        gi, gd = gi[rows], gd[rows]
        recall = float(np.mean([np.intersect1d(a, b).size / k for a, b in zip(gi, ri)]))
        check = {"rows": int(q), "recall_at_k": recall,
                 "rows_bit_exact": int(np.sum(np.all(gi == ri, axis=1) & np.all(gd == rd, axis=1)))}
    # connectivities + Leiden on a subsample of m cells (cost ~ linear in cells at fixed k)
    m = min(n, 20_000)
    sub = np.ascontiguousarray(X[rng.choice(n, m, replace=False)])
    si, _, sd = knn.exact_knn_c(sub, k, threads=threads)
    t = time.perf_counter()
    P, _, _ = fuzzy.fuzzy_simplicial_set(si, sd)
    t_conn = (time.perf_counter() - t) * n / m
    t = time.perf_counter()
    leiden.leiden(P, resolution=1.0, random_state=0, n_iterations=2)
    t_leiden = (time.perf_counter() - t) * n / m
    total = t_knn + t_conn + t_leiden
    total1 = t_knn1 + t_conn + t_leiden
    return {"value": n / total, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": (f"kNN: {q} query rows x all {n} candidates (C oracle, {threads} threads), "
                       f"x{n / q:.0f}; connectivities + Leiden (NumPy oracle, one thread, n_iterations=2) "
                       f"on a {m}-cell subsample, x{n / m:.0f}; all extrapolated linearly to {n} cells"),
            "stage_s": {"knn": t_knn, "connectivities": t_conn, "leiden": t_leiden},
            "single_thread": {"value": n / total1, "cores": 1, "knn_s": t_knn1,
                              "sample": f"kNN: {q1} query rows on one thread, x{n / q1:.0f}; "
                                        "connectivities and Leiden as above"},
            "oracle_check": check}


def bf16_ranking_error(X, idx, rows):
    """Distance error of the BF16 ranking form against float64, on (row, neighbour) pairs:
    the device ranks |q^ - c~|^2 with q^ = bf16(q - mu) and c~ = hi + lo, hi = bf16(c - mu),
    lo = bf16((c - mu) - hi); this evaluates that squared distance exactly (float64) from the
    rounded operands and compares it with the float64 squared distance of the float32 rows.
    (The FP32 accumulation error is bounded separately by the certificate.)"""
    import numpy as np
    import torch
    Xd = torch.from_numpy(np.ascontiguousarray(X, np.float32))
    mu = Xd.double().mean(0).float()
    C = Xd - mu
    hi = C.to(torch.bfloat16).float()
    lo = (C - hi).to(torch.bfloat16).float()
    q_hat = hi[rows].double()
    nb = torch.from_numpy(np.asarray(idx[rows], np.int64))
    c_til = hi[nb].double() + lo[nb].double()
    approx = ((q_hat[:, None, :] - c_til) ** 2).sum(-1)
    X64 = Xd.double()
    exact = ((X64[rows][:, None, :] - X64[nb]) ** 2).sum(-1)
    mask = exact > 0
    rel = ((approx - exact).abs()[mask] / exact[mask]).numpy()
    return {"pairs": int(mask.sum()), "max_rel_d2": float(rel.max()), "median_rel_d2": float(np.median(rel)),
            "max_abs_d": float((approx.sqrt() - exact.sqrt()).abs().max())}


def precision_report(k_c3, idx_c3, X_c3):
    """BF16 ranking-distance error against float64 per configuration (the device kNN's own
    neighbours of 500 sampled rows each): the 1M Zipf mixture, the 100k anisotropic mixture
    and the largest of the 256 batched samples."""
    import numpy as np
    from metalsinglecell import _hip, synthetic
    rng = np.random.default_rng(7)
    out = {"c3_1M_zipf": bf16_ranking_error(X_c3, idx_c3, np.sort(rng.choice(X_c3.shape[0], 500, replace=False)))}
    X2, _ = synthetic.gaussian_mixture_anisotropic()
    i2, _ = _hip.knn_exact(X2, 15, precision="bf16")
    out["c2_100k_aniso"] = bf16_ranking_error(X2, i2.cpu().numpy(), np.sort(rng.choice(X2.shape[0], 500, replace=False)))
    S = [x for x, _ in synthetic.batched_samples()]
    big = max(S, key=lambda x: x.shape[0])
    i4, _ = _hip.knn_exact(big, 15, precision="bf16")
    out["c4_largest_sample"] = bf16_ranking_error(big, i4.cpu().numpy(),
                                                  np.sort(rng.choice(big.shape[0], 500, replace=False)))
    return out


def end_to_end(X, k):
    """Host embedding -> AnnData slots: pp.neighbors (upload, kNN, connectivities, CSR
    download into obsp) then tl.leiden(backend="gpu") (CSR upload, Leiden, labels into a
    natsorted Categorical in obs); wall time of one call each after the warm device path."""
    import numpy as np
    from metalsinglecell import pp, tl
    from metalsinglecell._anndata import AnnData
    a = AnnData(n_obs=X.shape[0])          # no variables: .X is unused with use_rep
    a.obsm["X_pca"] = X
    t0 = time.perf_counter()
    pp.neighbors(a, n_neighbors=k, use_rep="X_pca")
    t1 = time.perf_counter()
    tl.leiden(a, backend="gpu")
    t2 = time.perf_counter()
    n = X.shape[0]
    return {"value": n / (t2 - t0), "unit": "cells/s", "neighbors_s": t1 - t0, "leiden_s": t2 - t1,
            "n_clusters": int(len(a.obs["leiden"].cat.categories)),
            "includes": "host->device upload of X, device kNN + connectivities, CSR download into "
                        "obsp['distances'/'connectivities'], canonical CSR upload, Leiden, labels "
                        "download, natsorted Categorical in obs['leiden']"}


def batched(reps=3, workers=8, per_sample=True):
    """256 samples, 500-60,000 cells log-uniform, d = 30, k = 15: the kNN + connectivities stage
    as one HIP-graph replay (split count from the autotuner), then per-sample Leiden."""
    import numpy as np
    import torch
    from metalsinglecell import _hip, synthetic
    from metalsinglecell.batch import BatchGraph, leiden_samples
    S = [x for x, _ in synthetic.batched_samples()]
    X = np.concatenate(S)
    sizes = [s.shape[0] for s in S]
    t = time.perf_counter()
    bg = BatchGraph(sizes, X.shape[1], k=15, precision="bf16", tune_on=X)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    g_ms, l_ms = [], []
    bg.run(X)
    leiden_samples(bg)                            # warm the joint clustering path
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bg.run(X)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        leiden_samples(bg)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        g_ms.append(1e3 * (t1 - t0))
        l_ms.append(1e3 * (t2 - t1))
    per_sample_leiden_ms = None
    if per_sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        leiden_samples(bg, workers=workers, method="per_sample")   # the per-sample form, for comparison
        torch.cuda.synchronize()
        per_sample_leiden_ms = 1e3 * (time.perf_counter() - t0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for x in S:                                   # the same stage as per-sample calls
        i, d = _hip.knn_exact(x, 15, precision="bf16")
        _hip.fuzzy_connectivities(i, d)
    torch.cuda.synchronize()
    per_sample_ms = 1e3 * (time.perf_counter() - t0)
    total = float(np.median(g_ms) + np.median(l_ms))
    return {"samples": len(S), "cells": int(X.shape[0]), "dims": int(X.shape[1]), "k": 15,
            "splits": bg.splits, "list_len": bg.list_len or 16, "setup_s": setup_s, "graph_stage_ms": float(np.median(g_ms)),
            "graph_stage_per_sample_calls_ms": per_sample_ms, "leiden_ms": float(np.median(l_ms)),
            "leiden_method": "one block-diagonal run of all samples (msc_leiden_batched)",
            "leiden_per_sample_calls_ms": per_sample_leiden_ms, "leiden_workers_per_sample": workers,
            "cells_per_s": X.shape[0] / (total * 1e-3)}


def serving(duration_s, rates=(20.0,), genes=500, n_perms=100, max_slots=6, seed=4):
    """Poisson streams of spatial jobs (grid kNN k=6 + Moran's I, P permutations) through the
    continuous-batching server, one stream per offered rate; arrival -> completion latency
    percentiles per rate, and the highest offered rate whose p99 stays under 2 s."""
    import threading
    import numpy as np
    from metalsinglecell import synthetic
    from metalsinglecell.serve import Server, SpatialJob

    streams = {}
    n_max = 0
    for rate in rates:
        n_jobs = max(1, int(duration_s * rate))
        arrivals, sizes = synthetic.poisson_job_stream(n_jobs, rate=rate, seed=seed)
        streams[rate] = (arrivals, sizes)
        n_max = max(n_max, int(sizes.max()))
    t = time.perf_counter()
    pool = synthetic.nb_counts(n_max, genes, rng=np.random.default_rng(seed))
    log(f"serving: gene pool {n_max} x {genes} in {time.perf_counter() - t:.1f}s")
    per_rate = {}
    for rate in rates:
        arrivals, sizes = streams[rate]
        rng = np.random.default_rng(seed + 1)
        jobs = [SpatialJob(coords=synthetic.tissue_coordinates(int(n), rng=rng), counts=pool[: int(n)],
                           n_perms=n_perms, seed=i, job_id=i) for i, n in enumerate(sizes)]
        srv = Server(max_slots=max_slots)
        futs = []
        t0 = time.perf_counter()

        def feed():
            for j, at in zip(jobs, arrivals):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                j.arrival = time.perf_counter()
                futs.append(srv.submit(j))

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        errors = 0
        for f in futs:
            try:
                f.wait(600)
            except Exception:
                errors += 1
        t1 = time.perf_counter()
        srv.close()
        st = srv.stats(t0, t1)
        st.update({"offered_rate_jobs_per_s": rate, "errors": errors, "genes": genes,
                   "n_perms": n_perms, "n_neighs": 6, "max_slots": max_slots,
                   "mean_cells": float(sizes.mean()), "hbm_budget_gib": srv.hbm_budget / 2**30,
                   "stream_s": float(arrivals[-1])})
        log(f"serving at {rate} jobs/s: p50 {st['p50_s']:.3f} s, p99 {st['p99_s']:.3f} s, "
            f"{st['jobs_per_s']:.1f} jobs/s completed")
        per_rate[str(rate)] = st
    ok = [r for r in rates if per_rate[str(r)]["p99_s"] < 2.0 and per_rate[str(r)]["errors"] == 0
          and per_rate[str(r)]["rejected"] == 0]
    first = per_rate[str(rates[0])]
    return {**first, "by_rate": per_rate, "max_offered_rate_p99_under_2s": max(ok) if ok else None}


def _ev_ms(fn, reps=3):
    """Median HIP-event time (ms) of fn() on torch's current stream, after one warm-up call;
    returns (ms, last result)."""
    import numpy as np
    import torch
    out = fn()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out = fn()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts)), out


def extensions(Xd, k):
    """The callers either side of the path, each at a production size:
    Louvain on the 1M-cell graph; batch-balanced kNN; fixed-radius spatial graph; Geary's C
    beside Moran's I; the cluster co-occurrence histogram."""
    import numpy as np
    import torch
    from metalsinglecell import _hip, synthetic
    out = {}
    n, d = Xd.shape
    # Louvain on the same connectivity graph as the Leiden stage
    idx, dist = _hip.knn_exact(Xd, k, precision="bf16")
    (ip, ind, dat), _ = _hip.fuzzy_connectivities(idx, dist)
    del idx, dist
    ms, (lab, q, info) = _ev_ms(lambda: _hip.louvain(ip, ind, dat, n), reps=2)
    out["louvain"] = {"cells": n, "edges": int(ind.numel()), "ms": ms, "modularity": q,
                      "levels": len(info["levels"]), "n_clusters": int(lab.max().item()) + 1,
                      "cells_per_s": n / (ms * 1e-3)}
    log("louvain", out["louvain"])
    del ip, ind, dat, lab
    # batch-balanced kNN: 500k cells x 50 dims in 8 batches, 3 neighbours per batch
    nb_n, nb_b = min(n, 500_000), 8
    Xb = Xd[:nb_n].contiguous()
    codes = np.random.default_rng(7).integers(0, nb_b, nb_n)
    ms, (bi, _) = _ev_ms(lambda: _hip.bbknn(Xb, codes, nb_b, 3, precision="bf16"), reps=2)
    _, _, binfo = _hip.bbknn(Xb, codes, nb_b, 3, precision="bf16", return_info=True)   # untimed diagnostics
    fl = 2.0 * nb_n * nb_n * d
    out["bbknn"] = {"cells": nb_n, "batches": nb_b, "k_per_batch": 3, "width": int(bi.shape[1]), "ms": ms,
                    "cells_per_s": nb_n / (ms * 1e-3), "select_flops": fl,
                    "select_tflops_incl_finalize": fl / (ms * 1e-3) / 1e12,
                    "fallback_lists": binfo["n_fallback"]}
    log("bbknn", out["bbknn"])
    del Xb, bi
    # fixed-radius spatial graph: 1M cells of tissue, radius for ~15 neighbours
    rng = np.random.default_rng(11)
    xy = synthetic.tissue_coordinates(1_000_000, rng=rng).astype(np.float32)
    xyd = torch.from_numpy(xy).cuda()
    area = float(np.ptp(xy[:, 0]) * np.ptp(xy[:, 1]))
    r = float(np.sqrt(15.0 * area / (np.pi * xy.shape[0])))
    ms, (rip, rind, _) = _ev_ms(lambda: _hip.spatial_radius(xyd, r), reps=3)
    pairs = int(rind.numel())
    out["radius"] = {"cells": xy.shape[0], "radius": r, "pairs": pairs, "ms": ms,
                     "pairs_per_s": pairs / (ms * 1e-3), "cells_per_s": xy.shape[0] / (ms * 1e-3)}
    log("radius", out["radius"])
    del rip, rind
    # Geary's C beside Moran's I: 50k spots x 500 genes, kNN(6) graph, 100 device permutations
    ns, G, P = 50_000, 500, 100
    sxy = xyd[:ns].contiguous()
    si, _ = _hip.spatial_knn(sxy, 6)
    sip = torch.arange(0, 6 * ns + 1, 6, dtype=torch.int64, device="cuda")
    sw = torch.ones(6 * ns, dtype=torch.float64, device="cuda")
    Xg = torch.from_numpy(synthetic.nb_counts(ns, G, rng=rng).toarray().astype(np.float32)).cuda()
    ms_g, _ = _ev_ms(lambda: _hip.geary(Xg, sip, si.reshape(-1), sw, n_perms=P, seed=1), reps=2)
    ms_m, _ = _ev_ms(lambda: _hip.moran(Xg, sip, si.reshape(-1), sw, n_perms=P, seed=1), reps=2)
    gathered = float(P + 1) * ns * G * 4.0
    out["geary"] = {"cells": ns, "genes": G, "n_perms": P, "ms": ms_g, "moran_ms": ms_m,
                    "gathered_GBps": gathered / (ms_g * 1e-3) / 1e9}
    log("geary", out["geary"])
    del Xg, si, sip, sw
    # cluster co-occurrence: 50k cells, 10 clusters, 50 thresholds (all ordered pairs)
    from metalsinglecell import spatial as _sp
    cn = 50_000
    cxy = xy[:cn]
    labels = rng.integers(0, 10, cn)
    lo, hi = _sp.cooccur_thresholds(cxy.astype(np.float64))
    thr2 = (np.linspace(lo, hi, 50)[1:] ** 2).astype(np.float32)
    cxyd = torch.from_numpy(np.ascontiguousarray(cxy)).cuda()
    ms, _ = _ev_ms(lambda: _hip.co_occurrence_hist(cxyd, labels, 10, thr2), reps=2)
    out["co_occurrence"] = {"cells": cn, "clusters": 10, "thresholds": 50, "ms": ms,
                            "pairs_per_s": float(cn) * (cn - 1) / (ms * 1e-3)}
    log("co_occurrence", out["co_occurrence"])
    return out


def extensions_cpu(X, threads, ext):
    """CPU timings beside each extension, on bounded samples of the same workloads (the
    oracle restatements, kind "port"; the radius graph uses scipy's k-d tree, kind "library")."""
    import numpy as np
    from scipy.spatial import cKDTree
    from metalsinglecell import synthetic
    from oracle import cooccur, fuzzy, knn, leiden, moran
    rng = np.random.default_rng(0)
    n = X.shape[0]
    out = {}
    # Louvain: a 20k-cell subsample graph, extrapolated linearly to n cells
    m = min(n, 20_000)
    sub = np.ascontiguousarray(X[rng.choice(n, m, replace=False)])
    si, _, sd = knn.exact_knn_c(sub, 30, threads=threads)
    P, _, _ = fuzzy.fuzzy_simplicial_set(si, sd)
    t = time.perf_counter()
    leiden.louvain(P)
    out["louvain"] = {"ms": (time.perf_counter() - t) * 1e3 * n / m, "kind": "port", "cores": 1,
                      "sample": f"NumPy Louvain on a {m}-cell subsample graph (k=30), x{n / m:.0f}"}
    # bbknn: 96 query rows per batch against that batch (C oracle), extrapolated to n x n_b pairs
    nb_n = ext["bbknn"]["cells"]
    codes = np.random.default_rng(7).integers(0, 8, nb_n)
    Xb = X[:nb_n]
    tot = 0.0
    for b in range(8):
        Xs = np.ascontiguousarray(Xb[codes == b])
        rows = rng.choice(Xs.shape[0], 96, replace=False)
        t = time.perf_counter()
        knn.exact_knn_c(Xs, 3, rows=rows, threads=threads)
        tot += (time.perf_counter() - t) * nb_n / 96
    out["bbknn"] = {"ms": tot * 1e3, "kind": "port", "cores": threads,
                    "sample": f"96 queries x each of 8 batches (C oracle, {threads} threads), scaled to {nb_n} queries"}
    # radius: scipy k-d tree over the same 1M points
    xy = synthetic.tissue_coordinates(1_000_000, rng=np.random.default_rng(11)).astype(np.float32)
    t = time.perf_counter()
    tree = cKDTree(xy.astype(np.float64))
    tree.query_ball_point(xy.astype(np.float64), ext["radius"]["radius"], workers=threads)
    out["radius"] = {"ms": (time.perf_counter() - t) * 1e3, "kind": "library", "cores": threads,
                     "sample": "scipy cKDTree build + query_ball_point, all 1M points"}
    # Geary: the float64 edge-form oracle, observed + 2 permutations, scaled to 100 permutations
    import scipy.sparse as sp
    ns, G = ext["geary"]["cells"], ext["geary"]["genes"]
    sxy = xy[:ns].astype(np.float64)
    _, nn = cKDTree(sxy).query(sxy, 7, workers=threads)
    W = sp.csr_matrix((np.ones(ns * 6), nn[:, 1:].ravel(), np.arange(0, 6 * ns + 1, 6)), shape=(ns, ns))
    Xg = synthetic.nb_counts(ns, G, rng=np.random.default_rng(12)).toarray()
    perms = moran.permutations(ns, 2, 1)
    t = time.perf_counter()
    moran.geary_scatter(Xg, moran.row_normalize(W), perms)
    el = time.perf_counter() - t
    out["geary"] = {"ms": el * 1e3 * (1 + ext["geary"]["n_perms"]) / 3.0, "kind": "port", "cores": 1,
                    "sample": f"NumPy float64 Geary, observed + 2 permutations at {ns} x {G}, scaled to "
                              f"{ext['geary']['n_perms']} permutations"}
    # co-occurrence: 1000 query rows against all cells, scaled to every row
    cn = ext["co_occurrence"]["cells"]
    cxy = xy[:cn]
    lab = rng.integers(0, 10, cn)
    thr = cooccur.thresholds(cxy.astype(np.float64), 50)
    q = rng.choice(cn, 1000, replace=False)
    t = time.perf_counter()
    _qhist(cxy, q, lab, (thr[1:] ** 2).astype(np.float32))
    out["co_occurrence"] = {"ms": (time.perf_counter() - t) * 1e3 * cn / 1000, "kind": "port", "cores": 1,
                            "sample": f"NumPy pair histogram of 1000 query rows x {cn} cells, x{cn / 1000:.0f}"}
    return out


def _qhist(xy, q, lab, thr2):
    """The oracle's per-pair binning for a subset of query rows (timing only)."""
    import numpy as np
    Q = xy[q]
    acc = None
    for t in range(xy.shape[1]):
        diff = (xy[None, :, t] - Q[:, t, None]).astype(np.float32)
        sq = (diff * diff).astype(np.float32)
        acc = sq if acc is None else (acc + sq).astype(np.float32)
    r = np.searchsorted(thr2, acc, side="left")
    keep = r < thr2.size
    hist = np.zeros((10, 10, thr2.size), np.int64)
    qi, j = np.nonzero(keep)
    np.add.at(hist, (lab[q][qi], lab[j], r[qi, j]), 1)
    return hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-rates", default="20,40,60,80", help="offered job rates (jobs/s), comma-separated")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-per-sample", action="store_true",
                    help="skip the per-sample (8 host threads) Leiden comparison of the batched section")
    ap.add_argument("--no-extensions", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    from metalsinglecell import synthetic

    torch.cuda.set_device(0)
    if a.copy_probe_mib > 0:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters calibrate the
        # counters read for the select kernel in the same profiled process
        src = torch.ones(a.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {a.copy_probe_mib} MiB read + {a.copy_probe_mib} MiB written")
        del src, dst
    t = time.perf_counter()
    X, _ = synthetic.gaussian_mixture_zipf(n=a.n, seed=2)
    log(f"data {X.shape} in {time.perf_counter() - t:.1f}s")
    Xd = torch.from_numpy(X).cuda()
    torch.cuda.synchronize()
    n, d = X.shape

    ms_bf, recs_bf, lab_bf = timed(Xd, a.k, "bf16", a.steps, a.warmup)
    log(f"bf16 form: {ms_bf * 1e3:.1f} ms/step", recs_bf[-1])
    fp32 = None
    if not a.no_fp32:
        ms_fp, recs_fp, lab_fp = timed(Xd, a.k, "fp32", 3, 1)
        same = bool(torch.equal(lab_bf, lab_fp))
        fp32 = {"value": n / ms_fp, "ms_per_step": ms_fp * 1e3, "steps": 3, "warmup": 1,
                "identical_labels": same, "knn_s": recs_fp[-1]["knn_s"],
                "select_ms": float(np.mean([r["knn"]["select_ms"] for r in recs_fp])),
                "speedup_bf16_vs_fp32": ms_fp / ms_bf}
        log("fp32 form:", fp32)
        # the plain FP32 form without the pruned candidate ranges (every query scans every row)
        ms_pl, recs_pl, lab_pl = timed(Xd, a.k, "fp32", 3, 1, prune=False)
        fp32["full_scan"] = {"value": n / ms_pl, "ms_per_step": ms_pl * 1e3, "steps": 3, "warmup": 1,
                             "identical_labels": bool(torch.equal(lab_bf, lab_pl)),
                             "knn_s": recs_pl[-1]["knn_s"],
                             "select_ms": float(np.mean([r["knn"]["select_ms"] for r in recs_pl])),
                             "speedup_bf16_pruned_vs_this": ms_pl / ms_bf}
        log("fp32 full-scan form:", fp32["full_scan"])

    sel = [r["knn"]["select_ms"] for r in recs_bf]
    sel_ms = float(np.mean(sel))
    # algorithmic flops of the launch: 2 d per (query, candidate) pair it scans (the pruned
    # layout scans `scanned` of the n^2 pairs); the brute-force equivalent is 2 n^2 d
    scanned = float(recs_bf[-1]["knn"].get("scanned", 1.0))
    flops = 2.0 * scanned * n * n * d
    achieved = flops / (sel_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "knn_select_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roof_sel = {"bound": "mfma", "kernel": "select32_kernel<7> (BF16, 32x32x16 MFMA)", "achieved": achieved,
                "peak": PEAK["bf16"], "unit": "TFLOP/s", "frac": achieved / PEAK["bf16"],
                "traffic": traffic, "flops_per_launch": flops, "avg_launch_ms": sel_ms,
                "scanned_fraction": scanned,
                "brute_force_equivalent_tflops": 2.0 * n * n * d / (sel_ms * 1e-3) / 1e12}
    roof = move_roofline(Xd, a.k, sel_ms)

    from metalsinglecell import _hip
    gi, gd = _hip.knn_exact(Xd, a.k, precision="bf16")
    gi, gd = gi.cpu().numpy(), gd.cpu().numpy()
    cpu = None
    if not a.no_cpu:
        threads = len(os.sched_getaffinity(0))
        t = time.perf_counter()
        cpu = cpu_baseline(X, a.k, threads, gpu_idx=(gi, gd))
        log(f"cpu baseline in {time.perf_counter() - t:.1f}s", cpu)
    prec = None
    if not a.no_precision:
        t = time.perf_counter()
        prec = precision_report(a.k, gi, X)
        log(f"precision report in {time.perf_counter() - t:.1f}s", prec)
    e2e = None
    if not a.no_e2e:
        end_to_end(X[:20_000], a.k)          # warm the host path once
        t = time.perf_counter()
        e2e = end_to_end(X, a.k)
        log(f"end to end in {time.perf_counter() - t:.1f}s", e2e)

    batch = None
    if not a.no_batched:
        t = time.perf_counter()
        batch = batched(per_sample=not a.no_per_sample)
        log(f"batched in {time.perf_counter() - t:.1f}s", batch)

    ext = None
    if not a.no_extensions:
        t = time.perf_counter()
        ext = extensions(Xd, a.k)
        log(f"extensions in {time.perf_counter() - t:.1f}s")
        if not a.no_cpu:
            t = time.perf_counter()
            ecpu = extensions_cpu(X, len(os.sched_getaffinity(0)), ext)
            for key, v in ecpu.items():
                ext[key]["cpu_baseline"] = v
                ext[key]["gpu_vs_cpu"] = v["ms"] / ext[key]["ms"]
            log(f"extension cpu baselines in {time.perf_counter() - t:.1f}s", ecpu)

    serve = None
    if a.serve_seconds > 0:
        t = time.perf_counter()
        serve = serving(a.serve_seconds, rates=tuple(float(r) for r in a.serve_rates.split(",")))
        log(f"serving in {time.perf_counter() - t:.1f}s", serve)

    last = recs_bf[-1]
    out = {
        "metric": "cells/sec kNN+Leiden at 1M×50 k=30 (BF16+rerank vs FP32); serve p50/p99",
        "value": n / ms_bf, "unit": "cells/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_bf * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"gaussian_mixture_zipf n={n} d={d} k={a.k} k'=64 "
                               "BF16 MFMA + exact FP32 re-rank over partition-pruned candidate ranges; "
                               "Leiden res=1.0 seed=0 n_iter=2",
                   "cells": n, "dims": d, "k": a.k, "parallelism": "single GPU"},
        "stages_s": {"knn": last["knn_s"], "connectivities": last["conn_s"], "leiden": last["leiden_s"]},
        "knn_info": last["knn"], "modularity": last["q"], "n_clusters": last["n_clusters"],
        "fp32_form": fp32, "precision": prec, "end_to_end": e2e, "batched": batch, "serving": serve,
        "extensions": ext, "roofline": roof, "roofline_select": roof_sel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
