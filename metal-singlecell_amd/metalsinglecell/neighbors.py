"""Neighbour graphs on the PCA embedding and on spatial coordinates (MI355X HIP kernels).

Mirrors the reference's ``metalsinglecell.neighbors`` seams:

* ``_knn(X, n_neighbors, ...)`` -> ``(indices, dists)`` (src/metalsinglecell/neighbors.py:698-724),
  self-inclusive, self in column 0, ascending.  Here it is always EXACT (the reference is
  exact below 30k cells and approximate NNDescent above): tiled MFMA distances with a fused
  top-k, an exact float32 re-rank and a per-row certificate (csrc/knn.hip).
* ``neighbors(X, n_neighbors, random_state)`` -> ``(distances, connectivities)`` scipy CSR
  (neighbors.py:727-760), connectivities from the fuzzy simplicial set (csrc/fuzzy.hip).
* ``_knn_grid(coords, k)`` -> ``(indices, dists)`` self-inclusive exact spatial kNN
  (neighbors.py:333-417) on the device grid (csrc/grid.hip).
* ``_radius_grid(coords, radius, min_radius)`` -> CSR of all pairs within the radius
  (neighbors.py:255-330), count and fill passes on the device grid (csrc/grid.hip).
* ``bbknn(X_pca, batch, ...)`` -> ``(distances, connectivities, n_neighbors)``
  (neighbors.py:491-559 with ``_trim_graph`` :465-488): per-batch exact kNN in one select
  launch (one candidate split per batch), connectivities and trimming on the device.

All computation is on the GPU; without one these functions raise.
"""

from __future__ import annotations

import numpy as np

from . import _hip

DEFAULT_PRECISION = "bf16"


def _csr_to_scipy(indptr, indices, data, n):
    """scipy CSR of a device CSR from the kernels here, whose rows hold sorted, distinct columns
    (so the matrix is marked canonical and later users skip their O(nnz) check)."""
    import scipy.sparse as sp
    m = sp.csr_matrix((_hip.to_host(data), _hip.to_host(indices), _hip.to_host(indptr)), shape=(n, n))
    m.has_canonical_format = True
    return m


def knn_device(X, n_neighbors, precision=DEFAULT_PRECISION):
    """Device tensors ``(idx int32 (n,k), dist float32 (n,k))`` — no host copies."""
    n = X.shape[0]
    if not 1 <= n_neighbors <= n:
        raise ValueError(f"n_neighbors={n_neighbors} must be in [1, n_obs={n}]")
    return _hip.knn_exact(X, int(n_neighbors), precision=precision)


def _knn(X_pca, n_neighbors, random_state=0, approx=None, backend="exact",
         precision=DEFAULT_PRECISION):
    """Exact kNN ``(indices int64 (n,k), dists float32 (n,k))``, self first.

    ``approx`` and ``backend`` are accepted for signature compatibility with the reference;
    ``approx=True`` and the NNDescent / IVF / pynndescent engines do not exist here — every
    size takes the exact path.  ``random_state`` does not affect exact results.
    """
    if approx:
        raise NotImplementedError("approximate kNN is not provided: this path is exact at every size")
    if backend not in ("exact", "nndescent"):
        raise ValueError(f"unknown kNN backend {backend!r} (exact)")
    idx, dist = knn_device(X_pca, n_neighbors, precision)
    return idx.cpu().numpy().astype(np.int64), dist.cpu().numpy()


def neighbors_device(X, n_neighbors=15, precision=DEFAULT_PRECISION):
    """kNN, distances CSR and connectivities CSR, all as device tensors.

    Returns a dict with ``idx``, ``dist`` (n x k), ``distances`` and ``connectivities`` as
    ``(indptr int64, indices int32, data float32)`` tuples, and ``sigma``/``rho``.
    """
    idx, dist = knn_device(X, n_neighbors, precision)
    dcsr = _hip.knn_distances_csr(idx, dist)
    conn, (sigma, rho) = _hip.fuzzy_connectivities(idx, dist)
    return {"idx": idx, "dist": dist, "distances": dcsr, "connectivities": conn,
            "sigma": sigma, "rho": rho}


def neighbors(X_pca, n_neighbors=15, random_state=0, approx=None, precision=DEFAULT_PRECISION):
    """``(distances, connectivities)`` scipy CSR matrices (neighbors.py:727-760 contract).

    ``distances`` holds k-1 entries per row (self dropped), columns ascending, explicit zeros
    kept for coincident points; ``connectivities`` is the symmetric fuzzy simplicial set.
    """
    if approx:
        raise NotImplementedError("approximate kNN is not provided: this path is exact at every size")
    X = np.asarray(X_pca, dtype=np.float32) if not hasattr(X_pca, "device") else X_pca
    n = X.shape[0]
    out = neighbors_device(X, n_neighbors, precision)
    return _csr_to_scipy(*out["distances"], n), _csr_to_scipy(*out["connectivities"], n)


def _knn_grid(coords, k, min_n=None):
    """Exact self-inclusive spatial kNN ``(indices int64 (n,k), dists float32 (n,k))``.

    Column 0 is the point itself at distance 0; the rest are the k-1 nearest other points
    ordered by (float32 distance, index).  ``min_n`` is accepted for signature compatibility.
    """
    X = np.ascontiguousarray(coords, dtype=np.float32)
    n = X.shape[0]
    if k == 1:
        return np.arange(n)[:, None].astype(np.int64), np.zeros((n, 1), np.float32)
    idx, dist = _hip.spatial_knn(X, int(k) - 1)
    idx = np.c_[np.arange(n), idx.cpu().numpy()].astype(np.int64)
    dist = np.c_[np.zeros(n, np.float32), dist.cpu().numpy()]
    return idx, dist


def _radius_grid(coords, radius, min_radius=0.0):
    """All neighbours within ``radius`` (and at least ``min_radius``), self excluded:
    ``(indptr int64, indices int32, distances float32)`` of an (n, n) CSR with columns
    ascending (neighbors.py:255-330 returns the same pairs in cell-walk order)."""
    X = np.ascontiguousarray(coords, dtype=np.float32)
    n = X.shape[0]
    if n == 0:
        return np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32)
    if X.ndim != 2 or X.shape[1] not in (2, 3):
        raise ValueError(f"radius search expects 2-D or 3-D coordinates, got {X.shape[-1]}-D")
    ip, ind, dist = _hip.spatial_radius(X, float(radius), float(min_radius))
    return ip.cpu().numpy(), ind.cpu().numpy(), dist.cpu().numpy()


def _batch_codes(batch):
    """Codes 0..B-1 of the present categories in ``np.unique`` order (neighbors.py:510-511)."""
    cats, codes = np.unique(np.asarray(batch), return_inverse=True)
    return codes.astype(np.int32).ravel(), int(cats.size)


def bbknn_device(X, batch, neighbors_within_batch=3, trim=None, precision=DEFAULT_PRECISION):
    """Batch-balanced kNN graph as device tensors: dict with ``idx``/``dist`` (n x K),
    ``distances`` and ``connectivities`` CSR tuples (connectivities trimmed), and ``K``."""
    codes, nb = _batch_codes(batch)
    idx, dist = _hip.bbknn(X, codes, nb, int(neighbors_within_batch), precision=precision)
    K = int(idx.shape[1])
    dcsr = _hip.bbknn_distances_csr(idx, dist)
    conn, _ = _hip.fuzzy_connectivities(idx, dist)
    t = 10 * K if trim is None else int(trim)
    if t > 0:
        conn = _hip.trim_graph(*conn, idx.shape[0], t)
    return {"idx": idx, "dist": dist, "distances": dcsr, "connectivities": conn, "K": K}


def bbknn(X_pca, batch, neighbors_within_batch: int = 3, n_pcs: int | None = 50, trim: int | None = None,
          random_state: int = 0, precision=DEFAULT_PRECISION):
    """Batch-balanced kNN (neighbors.py:491-559): ``(distances, connectivities, n_neighbors)``.

    For every cell the ``min(neighbors_within_batch, n_b)`` exact nearest cells of each batch
    b (the cell itself first in its own batch), merged and sorted by distance (ties keep the
    batch order), then one fuzzy simplicial set over the K combined neighbours, trimmed to
    each cell's ``10 K`` (or ``trim``) strongest connectivities.  ``random_state`` does not
    affect the exact results."""
    X = np.asarray(X_pca, dtype=np.float32)
    if n_pcs is not None:
        X = X[:, :n_pcs]
    n = X.shape[0]
    if np.asarray(batch).shape[0] != n:
        raise ValueError("batch must have one entry per cell")
    out = bbknn_device(np.ascontiguousarray(X), batch, neighbors_within_batch, trim, precision)
    return _csr_to_scipy(*out["distances"], n), _csr_to_scipy(*out["connectivities"], n), out["K"]
