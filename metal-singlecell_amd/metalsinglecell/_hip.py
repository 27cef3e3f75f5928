"""ctypes binding of libmsc_hip (the gfx950 neighbor-graph kernels).

PyTorch supplies device memory and the current stream; every computation runs in the HIP
library.  There is no CPU fallback: without a visible ROCm device or without the built
library, every entry point raises.

``torch`` is imported before the library is loaded so that the library binds to the same
HIP runtime instance as torch (both resolve ``libamdhip64.so.7``); device pointers and
streams are then shared.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_LIB = None
_LOCK = threading.Lock()
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libmsc_hip.so")

PRECISION = {"fp32": 0, "bf16": 1, "bf16_rerank": 1}

c_i64, c_int, c_vp, c_sz, c_dbl = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double


class LeidenLevel(ctypes.Structure):
    _fields_ = [("n", c_i64), ("n_edges", c_i64), ("move_passes", ctypes.c_int32),
                ("refine_passes", ctypes.c_int32), ("n_comms", c_i64), ("n_refined", c_i64),
                ("modularity", c_dbl), ("move_delta_passes", ctypes.c_int32),
                ("refine_delta_passes", ctypes.c_int32)]


class LeidenInfo(ctypes.Structure):
    _fields_ = [("n_levels", c_int), ("scale_exponent", c_int), ("twom_fixed", c_i64),
                ("levels", LeidenLevel * 64), ("time_moves", ctypes.c_int32), ("move_launches", ctypes.c_int32),
                ("move_ms", c_dbl), ("move_model_bytes", c_dbl)]


class SpatialInfo(ctypes.Structure):
    _fields_ = [("max_shell", ctypes.c_int32), ("grid_x", ctypes.c_int32), ("grid_y", ctypes.c_int32),
                ("grid_z", ctypes.c_int32), ("cell", ctypes.c_float)]


class KnnInfo(ctypes.Structure):
    _fields_ = [("timing", c_int), ("n_splits", c_int), ("list_len", c_int),
                ("n_second_pass", c_i64), ("n_fallback", c_i64), ("prep_ms", c_dbl),
                ("select_ms", c_dbl), ("finalize_ms", c_dbl), ("second_pass_ms", c_dbl),
                ("fallback_ms", c_dbl), ("prune", c_int), ("n_partitions", c_int), ("n_groups", c_int),
                ("n_ranges", c_int), ("scanned", c_dbl), ("partition_ms", c_dbl)]


PRUNE = {None: 0, "auto": 0, False: 1, "off": 1, True: 2, "on": 2}


# name -> (restype, argtypes); the declarations of include/msc_hip.h
SIGNATURES = {
    "msc_version": (c_int, []),
    "msc_last_error": (ctypes.c_char_p, []),
    "msc_knn_workspace_size": (c_int, [c_i64, c_int, c_int, c_int, ctypes.POINTER(c_sz)]),
    "msc_knn_exact": (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_sz, c_vp,
                              ctypes.POINTER(KnnInfo)]),
    "msc_knn_batched_workspace_size": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int,
                                               ctypes.POINTER(c_sz)]),
    "msc_knn_batched_prepare": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_vp, c_sz, c_vp]),
    "msc_knn_batched_launch": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                       c_sz, c_vp]),
    "msc_knn_batched": (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_sz,
                                c_vp, ctypes.POINTER(KnnInfo)]),
    "msc_knn_distances_csr": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_vp]),
    "msc_bbknn_workspace_size": (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_int, ctypes.POINTER(c_sz)]),
    "msc_bbknn": (c_int, [c_vp, c_i64, c_int, c_vp, c_int, c_int, c_int, c_vp, c_vp, ctypes.POINTER(c_int),
                          c_vp, c_sz, c_vp, ctypes.POINTER(KnnInfo)]),
    "msc_bbknn_distances_workspace_size": (c_int, [c_i64, ctypes.POINTER(c_sz)]),
    "msc_bbknn_distances_csr": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "msc_trim_graph_workspace_size": (c_int, [c_i64, ctypes.POINTER(c_sz)]),
    "msc_trim_graph": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, ctypes.POINTER(c_i64), c_vp,
                               c_sz, c_vp]),
    "msc_fuzzy_workspace_size": (c_int, [c_i64, c_int, ctypes.POINTER(c_sz)]),
    "msc_fuzzy_connectivities": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_i64, c_vp,
                                         c_vp, ctypes.POINTER(c_i64), c_vp, c_sz, c_vp]),
    "msc_fuzzy_batched_workspace_size": (c_int, [c_i64, c_int, c_int, ctypes.POINTER(c_sz)]),
    "msc_fuzzy_connectivities_batched": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp, c_int, c_vp, c_vp, c_vp,
                                                 c_i64, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "msc_fuzzy_connectivities_batched_noalloc": (c_int, [c_vp, c_vp, c_i64, c_int, c_vp, c_int, c_vp, c_vp,
                                                         c_vp, c_i64, c_vp, c_vp, c_vp]),
    "msc_spatial_knn_workspace_size": (c_int, [c_i64, c_int, c_int, ctypes.POINTER(c_sz)]),
    "msc_spatial_knn": (c_int, [c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp,
                                ctypes.POINTER(SpatialInfo), c_vp, c_sz, c_vp]),
    "msc_spatial_radius_workspace_size": (c_int, [c_i64, c_int, c_i64, ctypes.POINTER(c_sz)]),
    "msc_spatial_radius_count": (c_int, [c_vp, c_i64, c_int, c_dbl, c_dbl, c_vp, ctypes.POINTER(c_i64), c_vp, c_sz,
                                         c_vp]),
    "msc_spatial_radius_fill": (c_int, [c_vp, c_i64, c_int, c_dbl, c_dbl, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp,
                                        c_sz, c_vp]),
    "msc_moran_workspace_size": (c_int, [c_i64, c_i64, c_i64, c_int, ctypes.POINTER(c_sz)]),
    "msc_moran": (c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_i64,
                          c_vp, c_vp, ctypes.POINTER(c_dbl), c_vp, c_sz, c_vp]),
    "msc_geary_workspace_size": (c_int, [c_i64, c_i64, c_i64, c_int, ctypes.POINTER(c_sz)]),
    "msc_geary": (c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_i64,
                          c_vp, c_vp, ctypes.POINTER(c_dbl), c_vp, c_sz, c_vp]),
    "msc_co_occurrence_workspace_size": (c_int, [c_vp, c_i64, c_int, c_int, c_int, ctypes.POINTER(c_sz)]),
    "msc_co_occurrence": (c_int, [c_vp, c_i64, c_int, c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_sz, c_vp]),
    "msc_csr_to_dense": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "msc_graph_moments_workspace_size": (c_int, [c_i64, c_i64, ctypes.POINTER(c_sz)]),
    "msc_graph_moments": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp]),
    "msc_leiden_workspace_size": (c_int, [c_i64, c_i64, ctypes.POINTER(c_sz)]),
    "msc_leiden": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_dbl, c_i64, c_int, c_dbl, c_vp, c_vp,
                           ctypes.POINTER(c_dbl), ctypes.POINTER(LeidenInfo), c_vp, c_sz, c_vp]),
    "msc_leiden_batched_workspace_size": (c_int, [c_i64, c_i64, c_int, ctypes.POINTER(c_sz)]),
    "msc_leiden_batched": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_int, c_dbl, c_i64, c_int, c_dbl,
                                   c_vp, c_vp, c_vp, c_sz, c_vp]),
    "msc_louvain_workspace_size": (c_int, [c_i64, c_i64, ctypes.POINTER(c_sz)]),
    "msc_louvain": (c_int, [c_vp, c_vp, c_vp, c_i64, c_i64, c_dbl, c_i64, c_int, c_dbl, c_vp,
                            ctypes.POINTER(c_dbl), ctypes.POINTER(LeidenInfo), c_vp, c_sz, c_vp]),
}


def load_library(path=LIB_PATH):
    """Load and type the library (no device needed; used by the symbol-export tests)."""
    global _LIB
    with _LOCK:
        if _LIB is not None:
            return _LIB
        if not os.path.exists(path):
            raise ImportError(f"libmsc_hip.so not found at {path}: build it with "
                              "`python metal-singlecell_amd/build_lib.py`")
        import torch  # noqa: F401  (bind to torch's HIP runtime)
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
        return lib


def check(rc):
    if rc != 0:
        msg = load_library().msc_last_error()
        raise RuntimeError(f"libmsc_hip error {rc}: {msg.decode() if msg else ''}")


def device():
    """The ROCm device every call runs on; raises when none is visible."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("metalsinglecell's neighbor-graph path needs a ROCm GPU "
                           "(torch.cuda.is_available() is False); there is no CPU fallback")
    load_library()
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def to_device(a, dtype):
    """numpy / torch -> contiguous device tensor of ``dtype``."""
    import torch
    dev = device()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype).contiguous()


def to_host(t):
    """Device tensor -> numpy array in ordinary host memory.  Large results go through a
    page-locked staging buffer (one DMA at full link speed, then a host copy): 400 MB in ~40 ms
    against ~100 ms for torch's pageable copy.  The result is not left in the page-locked
    buffer, which the CPU reads an order of magnitude slower here."""
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.numel() * t.element_size() < (1 << 20):
        return t.cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t)
    return h.numpy().copy()


def workspace(nbytes):
    import torch
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())


# ----------------------------------------------------------------------------------------
# (a) exact kNN
# ----------------------------------------------------------------------------------------
def knn_exact(X, k, precision="fp32", timing=False, return_info=False, prune=None):
    """Exact self-inclusive kNN on the device: ``(idx int32 (n,k), dist float32 (n,k))``.

    ``prune``: None / "auto" (partition-pruned candidate ranges from 65,536 rows on), False /
    "off" (every query scans every candidate) or True / "on"; the result is the same either way.
    """
    import torch
    lib = load_library()
    Xd = to_device(X, torch.float32)
    if Xd.dim() != 2:
        raise ValueError("X must be 2-D (cells x dims)")
    n, d = Xd.shape
    p = PRECISION[precision]
    nbytes = c_sz(0)
    check(lib.msc_knn_workspace_size(n, d, int(k), p, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    idx = torch.empty((n, k), dtype=torch.int32, device=Xd.device)
    dist = torch.empty((n, k), dtype=torch.float32, device=Xd.device)
    info = KnnInfo()
    info.timing = int(bool(timing))
    info.prune = PRUNE[prune]
    check(lib.msc_knn_exact(ptr(Xd), n, d, int(k), p, ptr(idx), ptr(dist), ptr(ws),
                            ws.numel(), stream_ptr(), ctypes.byref(info)))
    if return_info or timing:
        return idx, dist, {f: getattr(info, f) for f, _ in KnnInfo._fields_}
    return idx, dist


def knn_batched_workspace_bytes(offsets, d, k, precision="bf16", splits=1, list_len=0):
    lib = load_library()
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nbytes = c_sz(0)
    check(lib.msc_knn_batched_workspace_size(off.ctypes.data, off.size - 1, int(d), int(k), PRECISION[precision],
                                             int(splits), int(list_len), ctypes.byref(nbytes)))
    return nbytes.value


def knn_batched(X, offsets, k, precision="bf16", splits=1, list_len=0, return_info=False):
    """Block-diagonal exact kNN of concatenated samples: ``(idx int32 global, dist float32)``.

    ``list_len`` is the select tile's per-lane list length: 0 (default, 16), 16, or 8 (k <= 32
    only: half the list registers, more waves per CU).  Results do not depend on it."""
    import torch
    lib = load_library()
    Xd = to_device(X, torch.float32)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    S = off.size - 1
    d = Xd.shape[1]
    p = PRECISION[precision]
    ws = workspace(knn_batched_workspace_bytes(off, d, k, precision, splits, list_len))
    n = Xd.shape[0]
    idx = torch.empty((n, k), dtype=torch.int32, device=Xd.device)
    dist = torch.empty((n, k), dtype=torch.float32, device=Xd.device)
    info = KnnInfo()
    check(lib.msc_knn_batched(ptr(Xd), off.ctypes.data, S, d, int(k), p, int(splits), int(list_len), ptr(idx),
                              ptr(dist), ptr(ws), ws.numel(), stream_ptr(), ctypes.byref(info)))
    if return_info:
        return idx, dist, {f: getattr(info, f) for f, _ in KnnInfo._fields_}
    return idx, dist


def knn_distances_csr(idx, dist):
    """Device CSR arrays ``(indptr int64, indices int32, data float32)`` of obsp['distances']."""
    import torch
    lib = load_library()
    n, k = idx.shape
    indptr = torch.empty(n + 1, dtype=torch.int64, device=idx.device)
    m = n * (k - 1)
    indices = torch.empty(max(m, 1), dtype=torch.int32, device=idx.device)
    data = torch.empty(max(m, 1), dtype=torch.float32, device=idx.device)
    check(lib.msc_knn_distances_csr(ptr(idx), ptr(dist), n, k, ptr(indptr), ptr(indices),
                                    ptr(data), stream_ptr()))
    return indptr, indices[:m], data[:m]


def bbknn(X, batch_codes, n_batches, k, precision="bf16", return_info=False):
    """Batch-balanced kNN: ``(idx int32 (n,K), dist float32 (n,K))`` with K = sum_t min(k, n_t)."""
    import torch
    lib = load_library()
    Xd = to_device(X, torch.float32)
    if Xd.dim() != 2:
        raise ValueError("X must be 2-D (cells x dims)")
    n, d = Xd.shape
    codes = np.ascontiguousarray(batch_codes, dtype=np.int32)
    if codes.shape != (n,):
        raise ValueError("batch codes must have one entry per cell")
    p = PRECISION[precision]
    nbytes = c_sz(0)
    check(lib.msc_bbknn_workspace_size(codes.ctypes.data, n, int(n_batches), d, int(k), p, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    counts = np.bincount(codes, minlength=int(n_batches))
    K = int(np.minimum(counts, int(k)).sum())
    idx = torch.empty((n, K), dtype=torch.int32, device=Xd.device)
    dist = torch.empty((n, K), dtype=torch.float32, device=Xd.device)
    width = c_int(0)
    info = KnnInfo()
    check(lib.msc_bbknn(ptr(Xd), n, d, codes.ctypes.data, int(n_batches), int(k), p, ptr(idx), ptr(dist),
                        ctypes.byref(width), ptr(ws), ws.numel(), stream_ptr(),
                        ctypes.byref(info) if return_info else None))
    assert width.value == K
    if return_info:
        return idx, dist, {f: getattr(info, f) for f, _ in KnnInfo._fields_}
    return idx, dist


def bbknn_distances_csr(idx, dist):
    """Device CSR ``(indptr, indices, data)`` of bbknn's obsp['distances'] (self and zeros dropped)."""
    import torch
    lib = load_library()
    n, K = idx.shape
    nbytes = c_sz(0)
    check(lib.msc_bbknn_distances_workspace_size(n, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    indptr = torch.empty(n + 1, dtype=torch.int64, device=idx.device)
    indices = torch.empty(max(n * K, 1), dtype=torch.int32, device=idx.device)
    data = torch.empty(max(n * K, 1), dtype=torch.float32, device=idx.device)
    check(lib.msc_bbknn_distances_csr(ptr(idx), ptr(dist), n, K, ptr(indptr), ptr(indices), ptr(data), ptr(ws),
                                      ws.numel(), stream_ptr()))
    m = int(indptr[n].item())
    return indptr, indices[:m], data[:m]


def trim_graph(indptr, indices, data, n, trim):
    """Device CSR of ``neighbors._trim_graph`` (keep entries >= both rows' trim-th largest)."""
    import torch
    lib = load_library()
    indptr = to_device(indptr, torch.int64)
    indices = to_device(indices, torch.int32)
    data = to_device(data, torch.float32)
    nbytes = c_sz(0)
    check(lib.msc_trim_graph_workspace_size(n, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    cap = max(int(indices.numel()), 1)
    o_ip = torch.empty(n + 1, dtype=torch.int64, device=indptr.device)
    o_ind = torch.empty(cap, dtype=torch.int32, device=indptr.device)
    o_dat = torch.empty(cap, dtype=torch.float32, device=indptr.device)
    nnz = c_i64(0)
    check(lib.msc_trim_graph(ptr(indptr), ptr(indices), ptr(data), n, int(trim), ptr(o_ip), ptr(o_ind), ptr(o_dat),
                             ctypes.byref(nnz), ptr(ws), ws.numel(), stream_ptr()))
    m = nnz.value
    return o_ip, o_ind[:m], o_dat[:m]


# ----------------------------------------------------------------------------------------
# (c) fuzzy simplicial set connectivities
# ----------------------------------------------------------------------------------------
def fuzzy_connectivities(idx, dist):
    """Device CSR ``(indptr int64, indices int32, data float32)`` plus ``(sigma, rho)``."""
    import torch
    lib = load_library()
    n, k = idx.shape
    nbytes = c_sz(0)
    check(lib.msc_fuzzy_workspace_size(n, k, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    cap = max(2 * n * (k - 1), 1)
    dev = idx.device
    indptr = torch.empty(n + 1, dtype=torch.int64, device=dev)
    indices = torch.empty(cap, dtype=torch.int32, device=dev)
    data = torch.empty(cap, dtype=torch.float32, device=dev)
    sigma = torch.empty(n, dtype=torch.float32, device=dev)
    rho = torch.empty(n, dtype=torch.float32, device=dev)
    nnz = c_i64(0)
    check(lib.msc_fuzzy_connectivities(ptr(idx), ptr(dist), n, k, ptr(indptr), ptr(indices),
                                       ptr(data), cap, ptr(sigma), ptr(rho), ctypes.byref(nnz),
                                       ptr(ws), ws.numel(), stream_ptr()))
    m = nnz.value
    return (indptr, indices[:m], data[:m]), (sigma, rho)


def fuzzy_batched_workspace_bytes(n, k, n_groups):
    lib = load_library()
    nbytes = c_sz(0)
    check(lib.msc_fuzzy_batched_workspace_size(int(n), int(k), int(n_groups), ctypes.byref(nbytes)))
    return nbytes.value


# ----------------------------------------------------------------------------------------
# (d) Leiden
# ----------------------------------------------------------------------------------------
def leiden(indptr, indices, data, n, resolution=1.0, random_state=0, n_iterations=2,
           commit_prob=0.9, init_labels=None, time_moves=False):
    """Leiden on a device CSR: ``(labels int32 device (n,), modularity, info dict)``.  With
    ``time_moves`` the info dict also holds the summed HIP-event time of the move-kernel
    launches (``move_ms``, ``move_launches``) and their modelled bytes (``move_model_bytes``:
    8 E_l + 24 n_l per launch)."""
    import torch
    lib = load_library()
    indptr = to_device(indptr, torch.int64)
    indices = to_device(indices, torch.int32)
    data = to_device(data, torch.float32)
    nnz = int(indices.numel())
    nbytes = c_sz(0)
    check(lib.msc_leiden_workspace_size(n, nnz, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    labels = torch.empty(n, dtype=torch.int32, device=indptr.device)
    init = None if init_labels is None else to_device(init_labels, torch.int32)
    q = c_dbl(0.0)
    info = LeidenInfo()
    info.time_moves = 1 if time_moves else 0
    check(lib.msc_leiden(ptr(indptr), ptr(indices), ptr(data), n, nnz, float(resolution),
                         int(random_state), int(n_iterations), float(commit_prob),
                         None if init is None else ptr(init), ptr(labels), ctypes.byref(q),
                         ctypes.byref(info), ptr(ws), ws.numel(), stream_ptr()))
    levels = [{f: getattr(info.levels[i], f) for f, _ in LeidenLevel._fields_}
              for i in range(info.n_levels)]
    out = {"scale_exponent": info.scale_exponent, "twom_fixed": info.twom_fixed, "levels": levels}
    if time_moves:
        out.update(move_ms=info.move_ms, move_launches=info.move_launches, move_model_bytes=info.move_model_bytes)
    return labels, q.value, out


def leiden_batched(indptr, indices, data, offsets, resolution=1.0, random_state=0, n_iterations=2,
                   commit_prob=0.9):
    """Leiden of every sample of a block-diagonal CSR (samples at row ``offsets[s]:offsets[s+1]``,
    global column ids) in one run: ``(labels int32 device (n,), modularity per sample (numpy))``;
    sample s's labels equal ``leiden`` on that sample alone."""
    import torch
    lib = load_library()
    indptr = to_device(indptr, torch.int64)
    indices = to_device(indices, torch.int32)
    data = to_device(data, torch.float32)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    S = off.size - 1
    n = int(off[-1])
    nnz = int(indptr[n].item())           # indices / data may be longer (capacity-sized)
    nbytes = c_sz(0)
    check(lib.msc_leiden_batched_workspace_size(n, nnz, S, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    labels = torch.empty(n, dtype=torch.int32, device=indptr.device)
    q = np.zeros(S, dtype=np.float64)
    check(lib.msc_leiden_batched(ptr(indptr), ptr(indices), ptr(data), n, nnz, off.ctypes.data, S,
                                 float(resolution), int(random_state), int(n_iterations), float(commit_prob),
                                 ptr(labels), q.ctypes.data, ptr(ws), ws.numel(), stream_ptr()))
    return labels, q


def louvain(indptr, indices, data, n, resolution=1.0, random_state=0, max_levels=20, commit_prob=0.9):
    """Multilevel Louvain on a device CSR: ``(labels int32 device (n,), modularity, info dict)``."""
    import torch
    lib = load_library()
    indptr = to_device(indptr, torch.int64)
    indices = to_device(indices, torch.int32)
    data = to_device(data, torch.float32)
    nnz = int(indices.numel())
    nbytes = c_sz(0)
    check(lib.msc_louvain_workspace_size(n, nnz, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    labels = torch.empty(n, dtype=torch.int32, device=indptr.device)
    q = c_dbl(0.0)
    info = LeidenInfo()
    check(lib.msc_louvain(ptr(indptr), ptr(indices), ptr(data), n, nnz, float(resolution),
                          int(random_state), int(max_levels), float(commit_prob), ptr(labels),
                          ctypes.byref(q), ctypes.byref(info), ptr(ws), ws.numel(), stream_ptr()))
    levels = [{f: getattr(info.levels[i], f) for f, _ in LeidenLevel._fields_}
              for i in range(info.n_levels)]
    return labels, q.value, {"scale_exponent": info.scale_exponent,
                             "twom_fixed": info.twom_fixed, "levels": levels}


# ----------------------------------------------------------------------------------------
# (b) spatial grid kNN
# ----------------------------------------------------------------------------------------
def spatial_knn(coords, k, coords64=None, return_info=False):
    """Exact spatial kNN (self excluded): ``(idx int32, dist float32[, len64 float64])``."""
    import torch
    lib = load_library()
    X = to_device(coords, torch.float32)
    n, dim = X.shape
    nbytes = c_sz(0)
    check(lib.msc_spatial_knn_workspace_size(n, dim, int(k), ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    idx = torch.empty((n, k), dtype=torch.int32, device=X.device)
    dist = torch.empty((n, k), dtype=torch.float32, device=X.device)
    X64 = None if coords64 is None else to_device(coords64, torch.float64)
    L64 = None if coords64 is None else torch.empty((n, k), dtype=torch.float64, device=X.device)
    info = SpatialInfo()
    check(lib.msc_spatial_knn(ptr(X), n, dim, int(k), ptr(idx), ptr(dist),
                              None if X64 is None else ptr(X64), None if L64 is None else ptr(L64),
                              ctypes.byref(info) if return_info else None, ptr(ws), ws.numel(),
                              stream_ptr()))
    out = (idx, dist) if L64 is None else (idx, dist, L64)
    if return_info:
        return out + ({f: getattr(info, f) for f, _ in SpatialInfo._fields_},)
    return out


# ----------------------------------------------------------------------------------------
# (e) Moran's I
# ----------------------------------------------------------------------------------------
def spatial_radius(coords, r_max, r_min=0.0, coords64=None):
    """Fixed-radius neighbours (self excluded), canonical CSR on the device:
    ``(indptr int64, indices int32, dist float32[, len64 float64])``."""
    import torch
    lib = load_library()
    X = to_device(coords, torch.float32)
    n, dim = X.shape
    nbytes = c_sz(0)
    check(lib.msc_spatial_radius_workspace_size(n, dim, 0, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    indptr = torch.empty(n + 1, dtype=torch.int64, device=X.device)
    nnz = c_i64(0)
    check(lib.msc_spatial_radius_count(ptr(X), n, dim, float(r_min), float(r_max), ptr(indptr), ctypes.byref(nnz),
                                       ptr(ws), ws.numel(), stream_ptr()))
    m = nnz.value
    check(lib.msc_spatial_radius_workspace_size(n, dim, m, ctypes.byref(nbytes)))
    if nbytes.value > ws.numel():
        ws = workspace(nbytes.value)
    indices = torch.empty(max(m, 1), dtype=torch.int32, device=X.device)
    dist = torch.empty(max(m, 1), dtype=torch.float32, device=X.device)
    X64 = None if coords64 is None else to_device(coords64, torch.float64)
    L64 = None if coords64 is None else torch.empty(max(m, 1), dtype=torch.float64, device=X.device)
    check(lib.msc_spatial_radius_fill(ptr(X), n, dim, float(r_min), float(r_max), ptr(indptr), m, ptr(indices),
                                      ptr(dist), None if X64 is None else ptr(X64), None if L64 is None else ptr(L64),
                                      ptr(ws), ws.numel(), stream_ptr()))
    out = (indptr, indices[:m], dist[:m])
    return out if L64 is None else out + (L64[:m],)


def _autocorr(kind, X, w_indptr, w_indices, w_data, n_perms, perms, seed, transformation):
    import torch
    lib = load_library()
    Xd = to_device(X, torch.float32)
    n, G = Xd.shape
    ip = to_device(w_indptr, torch.int64)
    ind = to_device(w_indices, torch.int32)
    dat = to_device(w_data, torch.float64)
    nnz = int(ind.numel())
    P = int(n_perms or 0)
    pr = None
    if P and perms is not None:
        pr = to_device(perms, torch.int32)
        if tuple(pr.shape) != (P, n):
            raise ValueError(f"perms must have shape ({P}, {n})")
    nbytes = c_sz(0)
    check(getattr(lib, f"msc_{kind}_workspace_size")(n, G, nnz, P, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    stat = torch.empty(G, dtype=torch.float64, device=Xd.device)
    sims = torch.empty((max(P, 1), G), dtype=torch.float64, device=Xd.device)
    S0 = c_dbl(0.0)
    check(getattr(lib, f"msc_{kind}")(ptr(Xd), n, G, ptr(ip), ptr(ind), ptr(dat), nnz, int(bool(transformation)), P,
                                      None if pr is None else ptr(pr), int(seed), ptr(stat), ptr(sims),
                                      ctypes.byref(S0), ptr(ws), ws.numel(), stream_ptr()))
    return stat, (sims[:P] if P else None), S0.value


def moran(X, w_indptr, w_indices, w_data, n_perms=0, perms=None, seed=0, transformation=True):
    """Moran's I on the device: ``(I float64 (G,), sims float64 (P, G) or None, S0)``."""
    return _autocorr("moran", X, w_indptr, w_indices, w_data, n_perms, perms, seed, transformation)


def geary(X, w_indptr, w_indices, w_data, n_perms=0, perms=None, seed=0, transformation=True):
    """Geary's C on the device: ``(C float64 (G,), sims float64 (P, G) or None, S0)``."""
    return _autocorr("geary", X, w_indptr, w_indices, w_data, n_perms, perms, seed, transformation)


def co_occurrence_hist(coords, codes, n_codes, thr2):
    """Per-bin ordered-pair counts ``(n_codes, n_codes, len(thr2))`` int64 on the device."""
    import torch
    lib = load_library()
    X = to_device(coords, torch.float32)
    n, dim = X.shape
    c = np.ascontiguousarray(codes, dtype=np.int32)
    t = np.ascontiguousarray(thr2, dtype=np.float32)
    if c.shape != (n,):
        raise ValueError("codes must have one entry per row")
    L = int(t.size)
    nbytes = c_sz(0)
    check(lib.msc_co_occurrence_workspace_size(c.ctypes.data, n, dim, int(n_codes), L, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    out = torch.empty((int(n_codes), int(n_codes), L), dtype=torch.int64, device=X.device)
    check(lib.msc_co_occurrence(ptr(X), n, dim, c.ctypes.data, int(n_codes), t.ctypes.data, L, ptr(out), ptr(ws),
                                ws.numel(), stream_ptr()))
    return out


def csr_to_dense(indptr, indices, data, n_rows, n_cols):
    """Dense float32 device matrix from a (host or device) CSR."""
    import torch
    lib = load_library()
    ip = to_device(indptr, torch.int64)
    ind = to_device(indices, torch.int32)
    dat = to_device(data, torch.float32)
    out = torch.empty((n_rows, n_cols), dtype=torch.float32, device=ip.device)
    check(lib.msc_csr_to_dense(ptr(ip), ptr(ind), ptr(dat), n_rows, n_cols, ptr(out), stream_ptr()))
    return out


def graph_moments(w_indptr, w_indices, w_data):
    """``(s0, s1, s2)`` of a (device or host) CSR weight matrix, computed on the device."""
    import torch
    lib = load_library()
    ip = to_device(w_indptr, torch.int64)
    ind = to_device(w_indices, torch.int32)
    dat = to_device(w_data, torch.float64)
    n = ip.numel() - 1
    nnz = int(ind.numel())
    nbytes = c_sz(0)
    check(lib.msc_graph_moments_workspace_size(n, nnz, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value)
    out = (c_dbl * 3)()
    check(lib.msc_graph_moments(ptr(ip), ptr(ind), ptr(dat), n, nnz, ctypes.cast(out, c_vp),
                                ptr(ws), ws.numel(),
                                stream_ptr()))
    return float(out[0]), float(out[1]), float(out[2])
