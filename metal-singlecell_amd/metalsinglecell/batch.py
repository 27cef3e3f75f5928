"""Batched form: neighbour graphs and Leiden clusterings of many independent samples.

Samples of uneven cell counts are concatenated (rows of sample s at ``offsets[s]:offsets[s+1]``).

* Graph stage, one HIP graph: the exact kNN of every sample in ONE select launch
  (``msc_knn_batched_launch``: each query block scans only its own sample) followed by the
  UMAP connectivities of every sample in one pass over the concatenation
  (``msc_fuzzy_connectivities_batched``; block s of the block-diagonal result equals the
  single-sample result).  Both enqueue kernels only, so the stage is captured once into a
  HIP graph (``torch.cuda.CUDAGraph``) and later batches of the same shape replay it.
* Clustering stage: the Leiden of every sample in one run over the block-diagonal graph
  (``msc_leiden_batched``): one pass loop in lockstep for all samples, each sample with its own
  fixed-point scale, total weight and coin stream, frozen when its own multilevel loop ends,
  so that its labels equal a single-sample run.  (``method="per_sample"`` instead calls
  ``msc_leiden`` per sample from several host threads, one stream each.)

Results are identical to per-sample calls of the single-sample entry points, whatever the
tile shape (split count, list length) the autotuner picks, and identical on every graph replay.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _hip, autotune


class BatchGraph:
    """Captured kNN + connectivities stage for a fixed set of sample sizes."""

    def __init__(self, sizes, d, k=15, precision="bf16", splits=None, capture=True, tune_on=None, list_len=None):
        import torch
        lib = _hip.load_library()
        self.sizes = np.asarray(sizes, dtype=np.int64)
        if self.sizes.ndim != 1 or self.sizes.size == 0 or (self.sizes < k).any():
            raise ValueError(f"every sample needs at least k={k} cells")
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.d, self.k, self.precision = int(d), int(k), precision
        dev = _hip.device()
        # static input: `tune_on` when given, else zeros, so that the warm-up run below (and a
        # result read before the first run()) never computes on uninitialised memory
        self.X = torch.zeros((self.n, self.d), dtype=torch.float32, device=dev)
        if tune_on is not None:
            self.X.copy_(_hip.to_device(tune_on, torch.float32))
        if splits is None:
            # tile shape: the cached (splits, list_len) of this shape bucket, else tuned on
            # `tune_on`, else one split with the default list
            shape = autotune.lookup(self.offsets, d, k, precision)
            if shape is None and tune_on is not None:
                shape, _ = autotune.tune_batched_tile(self.X, self.offsets, k, precision)
            if shape is not None:
                splits, tuned_len = shape
                list_len = tuned_len if list_len is None else list_len
        self.splits = int(splits or 1)
        self.list_len = int(list_len or 0)
        S = self.sizes.size
        self.ws_knn = _hip.workspace(_hip.knn_batched_workspace_bytes(self.offsets, d, k, precision, self.splits,
                                                                      self.list_len))
        self.ws_fz = _hip.workspace(_hip.fuzzy_batched_workspace_bytes(self.n, k, S))
        self.goff = torch.from_numpy(self.offsets).to(dev)
        self.idx = torch.empty((self.n, k), dtype=torch.int32, device=dev)
        self.dist = torch.empty((self.n, k), dtype=torch.float32, device=dev)
        cap = max(2 * self.n * (k - 1), 1)
        self.indptr = torch.empty(self.n + 1, dtype=torch.int64, device=dev)
        self.indices = torch.empty(cap, dtype=torch.int32, device=dev)
        self.data = torch.empty(cap, dtype=torch.float32, device=dev)
        self.sigma = torch.empty(self.n, dtype=torch.float32, device=dev)
        self.rho = torch.empty(self.n, dtype=torch.float32, device=dev)
        self._cap = cap
        _hip.check(lib.msc_knn_batched_prepare(self.offsets.ctypes.data, S, self.d, self.k,
                                               _hip.PRECISION[precision], self.splits, self.list_len,
                                               _hip.ptr(self.ws_knn),
                                               self.ws_knn.numel(), _hip.stream_ptr()))
        self.graph = None
        if capture:
            # capture on a side stream (torch.cuda.graph's own); the first eager run warms up
            self._enqueue()
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._enqueue()
            self.graph = g

    def _enqueue(self):
        lib = _hip.load_library()
        st = _hip.stream_ptr()
        S = self.sizes.size
        _hip.check(lib.msc_knn_batched_launch(_hip.ptr(self.X), self.offsets.ctypes.data, S, self.d, self.k,
                                              _hip.PRECISION[self.precision], self.splits, self.list_len,
                                              _hip.ptr(self.idx),
                                              _hip.ptr(self.dist), _hip.ptr(self.ws_knn), self.ws_knn.numel(), st))
        _hip.check(lib.msc_fuzzy_connectivities_batched(
            _hip.ptr(self.idx), _hip.ptr(self.dist), self.n, self.k, _hip.ptr(self.goff), S, _hip.ptr(self.indptr),
            _hip.ptr(self.indices), _hip.ptr(self.data), self._cap, _hip.ptr(self.sigma), _hip.ptr(self.rho),
            _hip.ptr(self.ws_fz), self.ws_fz.numel(), st))

    def run(self, X):
        """kNN and connectivities of the concatenated samples ``X`` (n x d)."""
        import torch
        src = _hip.to_device(X, torch.float32)
        if tuple(src.shape) != (self.n, self.d):
            raise ValueError(f"X must be {self.n} x {self.d}")
        self.X.copy_(src)
        if self.graph is not None:
            self.graph.replay()
        else:
            self._enqueue()
        return self.idx, self.dist, (self.indptr, self.indices, self.data)


def leiden_samples(bg, resolution=1.0, random_state=0, n_iterations=2, workers=8, method="joint"):
    """Leiden of every sample of a BatchGraph's last result: ``(labels, modularities)``, a list of
    int32 label tensors (dense per sample) and the per-sample modularity."""
    import torch
    lo_hi = [(int(bg.offsets[s]), int(bg.offsets[s + 1])) for s in range(bg.sizes.size)]
    if method == "joint":
        lab, q = _hip.leiden_batched(bg.indptr, bg.indices, bg.data, bg.offsets, resolution=resolution,
                                     random_state=random_state, n_iterations=n_iterations)
        return [lab[a:b] for a, b in lo_hi], [float(x) for x in q]
    if method != "per_sample":
        raise ValueError(f"method must be 'joint' or 'per_sample', got {method!r}")
    ip_host = bg.indptr.cpu().numpy()      # one transfer: row starts of every sample
    dev = bg.indptr.device

    # one stream per worker thread, created here once (default priority) and reused by every
    # sample that thread runs
    import threading
    streams = [torch.cuda.Stream(device=dev) for _ in range(max(1, int(workers)))]
    slot = threading.local()
    taken = iter(range(len(streams)))
    lock = threading.Lock()

    def one(s):
        a, b = lo_hi[s]
        e0, e1 = int(ip_host[a]), int(ip_host[b])
        if not hasattr(slot, "stream"):
            with lock:
                slot.stream = streams[next(taken)]
        stream = slot.stream
        with torch.cuda.stream(stream):
            ip = bg.indptr[a:b + 1] - e0
            ind = bg.indices[e0:e1] - a
            dat = bg.data[e0:e1]
            lab, q, _ = _hip.leiden(ip, ind, dat, b - a, resolution=resolution, random_state=random_state,
                                    n_iterations=n_iterations)
            stream.synchronize()
        return lab, q

    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        res = list(ex.map(one, range(bg.sizes.size)))
    return [r[0] for r in res], [r[1] for r in res]


def neighbors_leiden_batched(samples, k=15, precision="bf16", resolution=1.0, random_state=0, n_iterations=2,
                             workers=8, splits=None, autotune_splits=True, method="joint"):
    """kNN + connectivities + Leiden for a list of (n_s x d) float32 samples.

    Returns ``(labels, modularities, graph)``: per-sample int32 label tensors and Q values,
    and the BatchGraph holding the block-diagonal kNN and connectivities."""
    sizes = [int(x.shape[0]) for x in samples]
    X = np.concatenate([np.asarray(x, dtype=np.float32) for x in samples])
    bg = BatchGraph(sizes, X.shape[1], k=k, precision=precision, splits=splits,
                    tune_on=X if autotune_splits else None)
    bg.run(X)
    labels, qs = leiden_samples(bg, resolution, random_state, n_iterations, workers, method=method)
    return labels, qs, bg
