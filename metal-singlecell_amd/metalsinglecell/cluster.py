"""Graph clustering (scanpy ``sc.tl.leiden`` backends), mirroring the reference's
``metalsinglecell.cluster.leiden`` (src/metalsinglecell/cluster.py:36-75).

* ``backend="gpu"``: the synchronous parallel Leiden of csrc/leiden.hip on the MI355X.
* ``backend="igraph"``: igraph ``community_leiden`` on the CPU, as in the reference; igraph
  is an optional dependency and its absence raises ImportError (never a silent switch).
"""

from __future__ import annotations

import numpy as np

from . import _hip


def _canonical_csr(connectivities):
    import scipy.sparse as sp
    csr = sp.csr_matrix(connectivities, dtype=np.float32, copy=True)
    csr.sum_duplicates()
    csr.sort_indices()
    if csr.shape[0] != csr.shape[1]:
        raise ValueError("connectivities must be square")
    return csr


def leiden_device(indptr, indices, data, n, resolution=1.0, random_state=0, n_iterations=2,
                  commit_prob=0.9):
    """Leiden on a device CSR; returns ``(labels int32 device tensor, modularity, info)``."""
    return _hip.leiden(indptr, indices, data, n, resolution=resolution,
                       random_state=random_state, n_iterations=n_iterations,
                       commit_prob=commit_prob)


def leiden(connectivities, resolution=1.0, random_state=0, n_iterations=2, backend="igraph",
           variant="sync", commit_prob=0.9):
    """Leiden labels (int64 numpy, dense 0..C-1) of a symmetric connectivity matrix."""
    if backend == "gpu":
        if variant != "sync":
            raise NotImplementedError(f"variant={variant!r}: only the synchronous variant ('sync') "
                                      "is implemented on the MI355X")
        if not 0.0 < commit_prob <= 1.0:
            raise ValueError("commit_prob must be in (0, 1]")
        csr = _canonical_csr(connectivities)
        lab, _, _ = leiden_device(csr.indptr, csr.indices, csr.data, csr.shape[0], resolution,
                                  random_state, n_iterations, commit_prob)
        return lab.cpu().numpy().astype(np.int64)
    if backend != "igraph":
        raise ValueError(f"unknown backend {backend!r} (gpu|igraph)")
    try:
        import igraph as ig
    except ImportError as exc:
        raise ImportError("backend='igraph' needs python-igraph, which is not installed; "
                          "use backend='gpu' for the MI355X Leiden") from exc
    import random as _random
    coo = connectivities.tocoo()
    upper = coo.row < coo.col
    edges = np.column_stack([coo.row[upper], coo.col[upper]])
    try:
        ig.set_random_number_generator(_random.Random(random_state))
        g = ig.Graph(n=connectivities.shape[0], edges=edges.tolist())
        g.es["weight"] = coo.data[upper].tolist()
        part = g.community_leiden(objective_function="modularity", weights="weight",
                                  resolution=resolution, n_iterations=n_iterations)
    finally:
        ig.set_random_number_generator(_random)
    return np.asarray(part.membership)
