"""Continuous-batching server for spatial jobs (grid kNN graph + Moran's I) on one MI355X.

A job is one tissue section: 2-D coordinates and a cells x genes count matrix.  Its work is
the ``gr.spatial_neighbors_knn`` graph (k = 6) and ``gr.spatial_autocorr`` Moran's I with a
permutation null — csrc/grid.hip and csrc/moran.hip.

Scheduling:
* Each job has an HBM footprint estimated before admission (``SpatialJob.footprint``).
* A dispatcher admits waiting jobs in arrival order while the estimated bytes of admitted
  jobs stay within ``hbm_budget`` and fewer than ``max_slots`` jobs are in flight; a smaller
  job may pass a waiting larger one unless that one has waited longer than ``max_bypass_s``.
  A job larger than the whole budget is rejected with an error result, never run into OOM.
* Every admitted job runs in its own worker thread on one of ``max_slots`` HIP streams (default
  priority) that the server creates once and hands from job to job, so jobs that arrive while
  others are on the GPU overlap with them on the device.  Reusing the streams also lets torch's
  caching allocator, which keeps freed blocks per stream, serve a job from the blocks of an
  earlier one instead of new device allocations.  The ctypes calls release the GIL, so host-side
  enqueueing of different jobs proceeds concurrently.
* Results are deterministic: a job's outputs are bit-identical alone and in any mix.

Latency of a job = completion time - arrival (submit) time; ``Server.stats()`` reports
p50 / p99 latency, jobs/s and cells/s.
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _hip


@dataclass
class SpatialJob:
    coords: np.ndarray                 # (n, 2) float32 or float64
    counts: object                     # scipy CSR (n, G) float32
    n_neighs: int = 6
    n_perms: int = 100
    seed: int = 0
    job_id: int = -1
    arrival: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(self.coords.shape[0])

    @property
    def n_genes(self):
        return int(self.counts.shape[1])

    def footprint(self):
        """Estimated device bytes while the job runs (inputs, graph, dense X, Y, partials)."""
        n, G, k, P = self.n, self.n_genes, self.n_neighs, max(self.n_perms, 1)
        nnz = int(self.counts.nnz)
        graph = n * (8 + 16 + 16 * k + 8 * k + 200)
        moran = 8 * n * G + min(1024, n // 256 + 1) * P * G * 8 + 16 * G * (P + 2)
        upload = 8 * nnz + 8 * (n + 1)
        return int(1.15 * (graph + moran + upload)) + (64 << 20)


def run_spatial_job(job: SpatialJob, permutations="device"):
    """Run one job on the current stream; returns host results."""
    import torch
    n, k = job.n, job.n_neighs
    dev = _hip.device()
    xy64 = torch.from_numpy(np.ascontiguousarray(job.coords, dtype=np.float64)).to(dev, non_blocking=False)
    xy32 = xy64.to(torch.float32)
    idx, dist, len64 = _hip.spatial_knn(xy32, k, coords64=xy64)
    indptr = torch.arange(0, (n + 1) * k, k, dtype=torch.int64, device=dev)
    w = torch.ones(n * k, dtype=torch.float64, device=dev)
    C = job.counts
    X = _hip.csr_to_dense(C.indptr, C.indices, C.data, n, C.shape[1])
    perms = None
    if permutations == "numpy" and job.n_perms:
        from .spatial import permutations_numpy
        perms = permutations_numpy(n, job.n_perms, job.seed)
    I, sims, S0 = _hip.moran(X, indptr, idx.view(-1), w, n_perms=job.n_perms, perms=perms,
                             seed=job.seed, transformation=True)
    wn = torch.full((n * k,), 1.0 / k, dtype=torch.float64, device=dev)
    s0, s1, s2 = _hip.graph_moments(indptr, idx.view(-1), wn)
    I_h = I.cpu().numpy()
    sims_h = None if sims is None else sims.cpu().numpy()
    return {"idx": idx.cpu().numpy(), "dist": dist.cpu().numpy(), "I": I_h, "sims": sims_h,
            "S0": S0, "moments": (s0, s1, s2)}


class _Future:
    def __init__(self, job):
        self.job = job
        self._ev = threading.Event()
        self.result = None
        self.error = None
        self.t_submit = job.arrival
        self.t_start = None
        self.t_done = None

    def done(self):
        return self._ev.is_set()

    def wait(self, timeout=None):
        self._ev.wait(timeout)
        if self.error is not None:
            raise self.error
        return self.result


class Server:
    def __init__(self, hbm_budget=None, max_slots=6, max_bypass_s=0.5, permutations="device"):
        import torch
        _hip.device()
        free, total = torch.cuda.mem_get_info()
        # memory torch's caching allocator holds but no tensor uses is free for these jobs too
        cached = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
        self.hbm_budget = int(hbm_budget if hbm_budget is not None else 0.7 * (free + cached))
        self.max_slots = int(max_slots)
        self.max_bypass_s = max_bypass_s
        self.permutations = permutations
        self._lock = threading.Condition()
        self._waiting = []
        self._in_use = 0
        self._active = 0
        self._done = []
        self._stop = False
        self._threads = []
        self._streams = [torch.cuda.Stream() for _ in range(self.max_slots)]  # free streams
        self._disp = threading.Thread(target=self._dispatch, daemon=True)
        self._disp.start()

    # -- client side ------------------------------------------------------------------------
    def submit(self, job: SpatialJob):
        job.arrival = time.perf_counter() if not job.arrival else job.arrival
        fut = _Future(job)
        with self._lock:
            if job.footprint() > self.hbm_budget:
                fut.error = MemoryError(f"job {job.job_id}: estimated {job.footprint() / 2**30:.2f} GiB "
                                        f"exceeds the HBM budget {self.hbm_budget / 2**30:.2f} GiB")
                fut.t_done = time.perf_counter()
                fut._ev.set()
                self._done.append(fut)
                return fut
            self._waiting.append(fut)
            self._lock.notify_all()
        return fut

    def close(self):
        with self._lock:
            self._stop = True
            self._lock.notify_all()
        self._disp.join()
        for t in self._threads:
            t.join()

    # -- scheduling -------------------------------------------------------------------------
    def _pick(self):
        now = time.perf_counter()
        free = self.hbm_budget - self._in_use
        for i, f in enumerate(self._waiting):
            if f.job.footprint() <= free:
                return i
            if i == 0 and now - f.t_submit > self.max_bypass_s:
                return None          # the head has waited long enough: no bypass
        return None

    def _dispatch(self):
        while True:
            with self._lock:
                while True:
                    if self._stop and not self._waiting:
                        return
                    i = self._pick() if self._active < self.max_slots else None
                    if i is not None:
                        break
                    self._lock.wait(0.005)
                fut = self._waiting.pop(i)
                self._in_use += fut.job.footprint()
                self._active += 1
                stream = self._streams.pop()  # _active <= max_slots streams are taken
            t = threading.Thread(target=self._work, args=(fut, stream), daemon=True)
            self._threads.append(t)
            t.start()

    def _work(self, fut, stream):
        import torch
        fut.t_start = time.perf_counter()
        try:
            with torch.cuda.stream(stream):
                fut.result = run_spatial_job(fut.job, self.permutations)
                stream.synchronize()
        except Exception as exc:  # reported through the future
            fut.error = exc
        fut.t_done = time.perf_counter()
        with self._lock:
            self._in_use -= fut.job.footprint()
            self._active -= 1
            self._streams.append(stream)
            self._done.append(fut)
            self._lock.notify_all()
        fut._ev.set()

    # -- metrics ----------------------------------------------------------------------------
    def stats(self, t0=None, t1=None):
        ok = [f for f in self._done if f.error is None]
        lat = np.array([f.t_done - f.t_submit for f in ok]) if ok else np.zeros(0)
        cells = sum(f.job.n for f in ok)
        start = t0 if t0 is not None else (min(f.t_submit for f in ok) if ok else 0.0)
        end = t1 if t1 is not None else (max(f.t_done for f in ok) if ok else 0.0)
        span = max(end - start, 1e-9)
        return {"jobs": len(ok), "rejected": len(self._done) - len(ok),
                "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
                "max_s": float(lat.max()) if lat.size else None,
                "jobs_per_s": len(ok) / span, "cells_per_s": cells / span, "span_s": span}
