"""Exact Euclidean kNN oracle — TEST INFRASTRUCTURE ONLY.

Restates the exact kNN contract of the reference:

* ``neighbors._exact_knn_rows`` (src/metalsinglecell/neighbors.py:420-462): rank by an exact
  distance computed from direct differences (never the Gram expansion, :423-434), order ties
  by ``(distance, index)`` (:457-461).
* ``neighbors._knn_gpu`` (neighbors.py:82-132): self-inclusive, self in column 0, rows
  ascending (:129-131); float32 distances.

The exact squared distance is the float32 sum over t = 0..d-1, in that order, of
``(x_t - y_t)^2`` with every subtraction, multiplication and addition rounded separately
(no fused multiply-add).  The order key is ``(j != i, d2, j)``, which pins the query itself
first even when duplicate points exist (the reference leaves that case to an unstable
argsort, neighbors.py:129).  Returned distances are ``sqrt(d2)`` rounded to float32.

``exact_knn`` is vectorised NumPy (fine to a few thousand points); ``exact_knn_c`` runs the
same definition from ``knn_ref.c`` with OpenMP for the larger cases and the CPU baseline.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def sqdist_rows(X, rows):
    """float32 squared distances, shape (len(rows), n), by the fixed sequential order."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    Q = X[np.asarray(rows)]
    acc = np.zeros((Q.shape[0], X.shape[0]), dtype=np.float32)
    for t in range(X.shape[1]):
        diff = X[None, :, t] - Q[:, t, None]
        acc += diff * diff
    return acc


def exact_knn(X, k, rows=None, block=256):
    """Exact self-inclusive kNN: ``(idx int64 (m, k), d2 float32 (m, k), dist float32)``."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n = X.shape[0]
    rows = np.arange(n) if rows is None else np.asarray(rows)
    if not 1 <= k <= n:
        raise ValueError("need 1 <= k <= n")
    out_i = np.empty((rows.size, k), np.int64)
    out_d = np.empty((rows.size, k), np.float32)
    cols = np.arange(n)
    for s in range(0, rows.size, block):
        rb = rows[s:s + block]
        D = sqdist_rows(X, rb)
        for r, i in enumerate(rb):
            d = D[r]
            kth = np.partition(d, k - 1)[k - 1]
            cand = np.flatnonzero(d <= kth)
            cand = np.union1d(cand, [i])
            self_flag = (cand != i).astype(np.int8)
            o = np.lexsort((cols[cand], d[cand], self_flag))[:k]
            out_i[s + r] = cand[o]
            out_d[s + r] = d[cand[o]]
    return out_i, out_d, np.sqrt(out_d).astype(np.float32)


def _lib():
    path = os.path.join(_HERE, "_build", "liboracle.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
    lib = ctypes.CDLL(path)
    lib.oracle_knn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                               ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_int]
    lib.oracle_knn.restype = ctypes.c_int
    return lib


def exact_knn_c(X, k, rows=None, threads=0):
    """Same definition as :func:`exact_knn`, in C (``knn_ref.c``), OpenMP over query rows."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    rows = np.ascontiguousarray(np.arange(n) if rows is None else rows, dtype=np.int64)
    out_i = np.empty((rows.size, k), np.int64)
    out_d = np.empty((rows.size, k), np.float32)
    rc = _lib().oracle_knn(X.ctypes.data, n, d, k, rows.ctypes.data, rows.size,
                           out_i.ctypes.data, out_d.ctypes.data, int(threads))
    if rc != 0:
        raise RuntimeError(f"oracle_knn failed ({rc})")
    return out_i, out_d, np.sqrt(out_d).astype(np.float32)


def distances_csr(idx, dist, n):
    """``obsp['distances']`` (neighbors.py:755-760): k-1 entries per row, self column dropped."""
    import scipy.sparse as sp
    k = idx.shape[1]
    rows = np.repeat(np.arange(n), k - 1)
    return sp.csr_matrix((dist[:, 1:].ravel(), (rows, idx[:, 1:].ravel())), shape=(n, n))
