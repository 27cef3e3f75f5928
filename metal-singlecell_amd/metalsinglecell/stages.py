"""Single-stage device workloads with their algorithmic work, for timing and profiling one
kernel family at a time:  ``python -m metalsinglecell.stages <stage> [--reps R]``.

Each stage builds its synthetic input, runs one warm-up call, then times ``reps`` calls with
HIP events on the current stream and prints one JSON line: the mean time, the stage's
algorithmic bytes (or flops) per call by the formulas below, and the achieved rate against
the MI355X peaks (8 TB/s HBM, 2.5 PFLOP/s dense BF16 MFMA).

Stages (sizes are the workloads' real sizes):
  select-c3   exact kNN of 1M x 50 (200 Zipf(1.5) clusters, seed 2), k = 30, BF16 + re-rank,
              with the partition-pruned candidate ranges and with the full scan; flops = 2 d per
              scanned (query, candidate) pair for the select kernel (its own event time)
  knn-c3      the whole exact kNN call of that workload (partition layout, prep, select, finalize,
              second pass, fallback), per-stage event times and the layout's statistics
  conn-c3     UMAP connectivities of that kNN; bytes = 8 n k + 8 E + 8 (n + 1)
  leiden-c3   Leiden (res 1.0, seed 0, 2 iterations) on those connectivities;
              bytes = sum over levels of (passes (8 E_l + 24 n_l) + 12 (E_l + E_l+1)); then one
              call with the move kernel's launches timed (move_kernel: 8 E_l + 24 n_l per launch)
  grid-2m     spatial kNN of 2M tissue cells (10 elliptical blobs), k = 6;
              bytes = n (24 + 8 k) + 4 (cells + 1), cells ~ n / max(2k, 8)
  moran-2m    Moran's I of 2M cells x 500 genes, kNN(6) weights, 100 device permutations;
              bytes = 4 n G (P + 3) + 8 nnz(W) + 4 n P
  batched-c4  256 samples of 500-60,000 cells x 30 (seed 3), k = 15: the captured kNN +
              connectivities graph stage at the autotuned tile shape (every candidate's time
              reported), the joint Leiden of every sample, and (unless --no-per-sample) the
              per-sample Leiden calls for comparison
"""

from __future__ import annotations

import argparse
import json
import sys

import numpy as np

HBM_PEAK = 8.0e12
BF16_PEAK = 2.5e15


def _events_ms(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(st)
        fn()
        b.record(st)
        b.synchronize()
        times.append(a.elapsed_time(b))
    return float(np.mean(times)), times


def _c3():
    import torch
    from . import _hip, synthetic
    X, _ = synthetic.gaussian_mixture_zipf(n=1_000_000, d=50, n_clusters=200, a=1.5, seed=2)
    return torch.from_numpy(X).to(_hip.device())


def stage_select_c3(reps):
    from . import _hip
    Xd = _c3()
    n, d = Xd.shape
    out = {}
    for mode in ("auto", "off"):
        infos = []
        for _ in range(reps + 1):
            _, _, info = _hip.knn_exact(Xd, 30, precision="bf16", timing=True, prune=mode)
            infos.append(info)
        sel = float(np.mean([i["select_ms"] for i in infos[1:]]))
        scanned = float(infos[-1]["scanned"])
        flops = 2.0 * scanned * n * n * d
        out["pruned" if mode == "auto" else "full_scan"] = {
            "ms": sel, "flops": flops, "tflops": flops / sel / 1e9, "frac": flops / (sel * 1e-3) / BF16_PEAK,
            "scanned": scanned, "brute_force_equivalent_tflops": 2.0 * n * n * d / sel / 1e9,
            "partition_ms": float(np.mean([i["partition_ms"] for i in infos[1:]])),
            "prep_ms": float(np.mean([i["prep_ms"] for i in infos[1:]])),
            "finalize_ms": float(np.mean([i["finalize_ms"] for i in infos[1:]])),
            "second_pass_ms": float(np.mean([i["second_pass_ms"] for i in infos[1:]])),
            "n_second_pass": infos[-1]["n_second_pass"], "n_fallback": infos[-1]["n_fallback"],
            "n_partitions": infos[-1]["n_partitions"], "n_groups": infos[-1]["n_groups"],
            "n_ranges": infos[-1]["n_ranges"]}
    return out


def stage_knn_c3(reps):
    """The whole exact kNN call at C3 (BF16, pruned candidate ranges): per-stage event times."""
    from . import _hip
    Xd = _c3()
    infos, walls = [], []
    import time
    import torch
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, _, info = _hip.knn_exact(Xd, 30, precision="bf16", timing=True)
        torch.cuda.synchronize()
        walls.append((time.perf_counter() - t0) * 1e3)
        infos.append(info)
    keys = ("partition_ms", "prep_ms", "select_ms", "finalize_ms", "second_pass_ms", "fallback_ms")
    out = {k: float(np.mean([i[k] for i in infos[1:]])) for k in keys}
    out["wall_ms"] = float(np.mean(walls[1:]))
    for k in ("scanned", "n_partitions", "n_groups", "n_ranges", "n_second_pass", "n_fallback"):
        out[k] = infos[-1][k]
    return out


def _c3_graph():
    from . import _hip
    Xd = _c3()
    idx, dist = _hip.knn_exact(Xd, 30, precision="bf16")
    return idx, dist


def stage_conn_c3(reps):
    from . import _hip
    idx, dist = _c3_graph()
    n, k = idx.shape
    out = {}

    def run():
        out["conn"] = _hip.fuzzy_connectivities(idx, dist)
    ms, _ = _events_ms(run, reps)
    E = int(out["conn"][0][1].numel())
    b = 8.0 * n * k + 8.0 * E + 8.0 * (n + 1)
    return {"ms": ms, "bytes": b, "gbps": b / ms / 1e6, "frac": b / (ms * 1e-3) / HBM_PEAK, "nnz": E}


def stage_leiden_c3(reps):
    from . import _hip
    idx, dist = _c3_graph()
    (ip, ind, dat), _ = _hip.fuzzy_connectivities(idx, dist)
    n = idx.shape[0]
    out = {}

    def run():
        out["r"] = _hip.leiden(ip, ind, dat, n, resolution=1.0, random_state=0, n_iterations=2)
    ms, _ = _events_ms(run, reps)
    levels = out["r"][2]["levels"]
    b = 0.0
    for i, lv in enumerate(levels):
        passes = lv["move_passes"] + lv["refine_passes"]
        b += passes * (8.0 * lv["n_edges"] + 24.0 * lv["n"])
        nxt = levels[i + 1]["n_edges"] if i + 1 < len(levels) and levels[i + 1]["n"] < lv["n"] else 0
        b += 12.0 * (lv["n_edges"] + nxt)
    passes = sum(lv["move_passes"] + lv["refine_passes"] for lv in levels)
    # one more call with every move-kernel launch timed by HIP events
    _, _, ti = _hip.leiden(ip, ind, dat, n, resolution=1.0, random_state=0, n_iterations=2, time_moves=True)
    mv = {"launches": ti["move_launches"], "ms": ti["move_ms"], "avg_launch_us": 1e3 * ti["move_ms"] / ti["move_launches"],
          "model_bytes": ti["move_model_bytes"],
          "gbps": ti["move_model_bytes"] / ti["move_ms"] / 1e6,
          "frac": ti["move_model_bytes"] / (ti["move_ms"] * 1e-3) / HBM_PEAK}
    return {"ms": ms, "bytes": b, "gbps": b / ms / 1e6, "frac": b / (ms * 1e-3) / HBM_PEAK, "passes": passes,
            "levels": len(levels), "modularity": out["r"][1], "move_kernel": mv}


def _tissue(n, genes=0):
    from . import synthetic
    rng = np.random.default_rng(4)
    xy = synthetic.tissue_coordinates(n, rng=rng)
    X = synthetic.nb_counts(n, genes, rng=rng) if genes else None
    return xy, X


def stage_grid_2m(reps):
    import torch
    from . import _hip
    n, k = 2_000_000, 6
    xy, _ = _tissue(n)
    dev = _hip.device()
    X = torch.from_numpy(np.ascontiguousarray(xy, dtype=np.float32)).to(dev)
    info = {}

    def run():
        info["r"] = _hip.spatial_knn(X, k, return_info=True)
    ms, _ = _events_ms(run, reps)
    gi = info["r"][2]
    cells = gi["grid_x"] * gi["grid_y"] * max(gi["grid_z"], 1)
    b = n * (24.0 + 8.0 * k) + 4.0 * (cells + 1)
    return {"ms": ms, "bytes": b, "gbps": b / ms / 1e6, "frac": b / (ms * 1e-3) / HBM_PEAK, "cells": cells,
            "max_shell": gi["max_shell"]}


def stage_moran_2m(reps):
    import torch
    from . import _hip
    n, G, P, k = 2_000_000, 500, 100, 6
    xy, C = _tissue(n, G)
    dev = _hip.device()
    idx, _ = _hip.spatial_knn(torch.from_numpy(xy.astype(np.float32)).to(dev), k)
    indptr = torch.arange(0, (n + 1) * k, k, dtype=torch.int64, device=dev)
    w = torch.ones(n * k, dtype=torch.float64, device=dev)
    X = _hip.csr_to_dense(C.indptr, C.indices, C.data, n, G)
    ind = idx.view(-1)

    def run():
        _hip.moran(X, indptr, ind, w, n_perms=P, perms=None, seed=0, transformation=True)
    ms, _ = _events_ms(run, reps)
    b = 4.0 * n * G * (P + 3) + 8.0 * n * k + 4.0 * n * P
    return {"ms": ms, "bytes": b, "gbps": b / ms / 1e6, "frac": b / (ms * 1e-3) / HBM_PEAK}


def stage_batched_c4(reps, per_sample=True):
    import time

    import torch
    from . import autotune, synthetic
    from .batch import BatchGraph, leiden_samples
    samples = [x for x, _ in synthetic.batched_samples(n_samples=256, lo=500, hi=60_000, d=30, seed=3)]
    sizes = [x.shape[0] for x in samples]
    X = np.concatenate(samples)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    shape = autotune.lookup(offsets, 30, 15, "bf16")
    Xd = torch.from_numpy(X).to(torch.device("cuda"))
    if shape is None:
        shape, _ = autotune.tune_batched_tile(Xd, offsets, 15, "bf16")
    tile_ms = autotune.load()[autotune.key_for(offsets, 30, 15, "bf16")]["ms"]
    bg = BatchGraph(sizes, 30, k=15, precision="bf16", splits=shape[0], list_len=shape[1])
    graph_ms, _ = _events_ms(lambda: bg.run(X), reps)
    out = {"per_sample": None}
    for method in (("joint", "per_sample") if per_sample else ("joint",)):
        leiden_samples(bg, method=method)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            leiden_samples(bg, method=method)
        torch.cuda.synchronize()
        out[method] = (time.perf_counter() - t) * 1e3 / reps
    leiden_ms = out["joint"]
    return {"splits": bg.splits, "list_len": bg.list_len or 16, "select_ms_by_tile": tile_ms,
            "graph_stage_ms": graph_ms, "leiden_ms": leiden_ms, "leiden_per_sample_calls_ms": out["per_sample"],
            "cells": int(X.shape[0]), "cells_per_s": X.shape[0] / ((graph_ms + leiden_ms) * 1e-3)}


STAGES = {"select-c3": stage_select_c3, "knn-c3": stage_knn_c3, "conn-c3": stage_conn_c3, "leiden-c3": stage_leiden_c3,
          "grid-2m": stage_grid_2m, "moran-2m": stage_moran_2m, "batched-c4": stage_batched_c4}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("stage", choices=sorted(STAGES))
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--no-per-sample", action="store_true", help="batched-c4: skip the per-sample Leiden calls")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (counter calibration)")
    a = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    if a.copy_probe_mib > 0:
        src = torch.ones(a.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print(f"[stage] {a.stage}", file=sys.stderr, flush=True)
    kw = {"per_sample": not a.no_per_sample} if a.stage == "batched-c4" else {}
    res = STAGES[a.stage](a.reps, **kw)
    print(json.dumps({"stage": a.stage, **res}), flush=True)


if __name__ == "__main__":
    main()
