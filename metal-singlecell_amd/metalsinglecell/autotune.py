"""Tile-shape autotuner for the batched neighbour-graph launch, with an on-disk cache.

The tile shape of the batched select launch (``msc_knn_batched_*(..., splits, list_len, ...)``)
has two knobs:

* ``splits``: with one split every query block of a sample scans the whole sample, so the
  largest samples form a long tail; more splits give more, shorter blocks at the cost of
  merging more partial lists;
* ``list_len``: keys each lane keeps per query, 16 (64 admitted candidates per query) or 8
  (32, for k <= 32): the short list halves the list registers and the upkeep of each
  admission, so more waves fit per CU, but its certificate fails more often and more rows
  take the exact fallback.

Results never depend on the choice (the exact re-rank and the certificate fix them), only
the time does.

Cache: a JSON object mapping keys ``"knn_batched|<precision>|d<d>|k<k>|max<2^a>|tot<2^b>"``
(power-of-two buckets of the largest sample and of the total cell count) to
``{"splits": s, "list_len": L, "ms": {"<s>x<L>": median ms}}``.  Location:
``$MSC_AUTOTUNE_CACHE`` or ``~/.cache/metalsinglecell/autotune.json``.
"""

from __future__ import annotations

import json
import os
import threading

import numpy as np

CANDIDATES = (1, 2, 4, 8)          # split counts
LIST_LENS = (16, 8)                 # per-lane list lengths (8 only for k <= 32)
_lock = threading.Lock()


def cache_path():
    return os.environ.get("MSC_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "metalsinglecell", "autotune.json"))


def _bucket(v):
    return 1 << max(0, int(np.ceil(np.log2(max(int(v), 1)))))


def key_for(offsets, d, k, precision):
    sizes = np.diff(np.asarray(offsets, dtype=np.int64))
    return (f"knn_batched|{precision}|d{int(d)}|k{int(k)}|max{_bucket(sizes.max())}"
            f"|tot{_bucket(sizes.sum())}")


def load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def store(key, entry):
    with _lock:
        table = load()
        table[key] = entry
        path = cache_path()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def lookup(offsets, d, k, precision):
    """The cached ``(splits, list_len)`` for this shape bucket, or None."""
    entry = load().get(key_for(offsets, d, k, precision))
    if entry is None:
        return None
    return int(entry["splits"]), int(entry.get("list_len", 0))


def candidates_for(k, splits=CANDIDATES, list_lens=LIST_LENS):
    """Tile shapes worth timing: every split count with the 16-key lists, and the 8-key lists
    where they hold k (k <= 32) with at least 4k admitted candidates per query over the splits
    (32 per split); with fewer, most rows fail the certificate and take the exact fallback."""
    return [(int(s), int(L)) for L in list_lens for s in splits
            if L == 16 or (k <= 4 * L and 4 * L * s >= 4 * k)]


def tune_batched_tile(X, offsets, k, precision="bf16", candidates=None, reps=3):
    """Time the batched select launch for each candidate tile shape ``(splits, list_len)`` on
    ``X`` (device rows of the concatenated samples), cache and return the fastest as
    ``((splits, list_len), {"<s>x<L>": ms})``."""
    import torch
    from . import _hip
    lib = _hip.load_library()
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    S = off.size - 1
    n, d = X.shape
    p = _hip.PRECISION[precision]
    idx = torch.empty((n, k), dtype=torch.int32, device=X.device)
    dist = torch.empty((n, k), dtype=torch.float32, device=X.device)
    timings = {}
    for s, L in (candidates if candidates is not None else candidates_for(k)):
        ws = _hip.workspace(_hip.knn_batched_workspace_bytes(off, d, k, precision, s, L))
        _hip.check(lib.msc_knn_batched_prepare(off.ctypes.data, S, d, int(k), p, int(s), int(L), _hip.ptr(ws),
                                               ws.numel(), _hip.stream_ptr()))
        ms = []
        for _ in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            _hip.check(lib.msc_knn_batched_launch(_hip.ptr(X), off.ctypes.data, S, d, int(k), p, int(s), int(L),
                                                  _hip.ptr(idx), _hip.ptr(dist), _hip.ptr(ws), ws.numel(),
                                                  _hip.stream_ptr()))
            e1.record()
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        timings[(s, L)] = float(np.median(ms[1:]))
        del ws
    best = min(timings, key=timings.get)
    table = {f"{s}x{L}": t for (s, L), t in timings.items()}
    store(key_for(off, d, k, precision), {"splits": int(best[0]), "list_len": int(best[1]), "ms": table})
    return best, table
