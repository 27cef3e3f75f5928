"""Graph clustering (scanpy ``sc.tl.leiden`` / ``sc.tl.louvain`` backends), mirroring the
reference's ``metalsinglecell.cluster.leiden`` (src/metalsinglecell/cluster.py:36-75) and the
backend switch of ``tl.louvain`` (src/metalsinglecell/tl.py:86-102).

* ``backend="gpu"``: the synchronous parallel Leiden / Louvain of csrc/leiden.hip on the MI355X.
* ``backend="igraph"``: igraph ``community_leiden`` / ``community_multilevel`` on the CPU, as
  in the reference; igraph
  is an optional dependency and its absence raises ImportError (never a silent switch).
"""

from __future__ import annotations

import numpy as np

from . import _hip


def _canonical_csr(connectivities):
    """float32 CSR with sorted, unique columns per row: the caller's arrays when they are that
    already (only read), else a canonical copy (the caller's matrix is never modified)."""
    import scipy.sparse as sp
    if sp.issparse(connectivities) and connectivities.format == "csr" and connectivities.dtype == np.float32:
        csr = connectivities  # keeps scipy's cached format flags
    else:
        csr = sp.csr_matrix(connectivities, dtype=np.float32)
    if csr.shape[0] != csr.shape[1]:
        raise ValueError("connectivities must be square")
    if not csr.has_canonical_format:
        csr = csr.copy()
        csr.sum_duplicates()   # sorts the columns of every row and sums repeated entries
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


def louvain_device(indptr, indices, data, n, resolution=1.0, random_state=0, max_levels=20, commit_prob=0.9):
    """Louvain on a device CSR; returns ``(labels int32 device tensor, modularity, info)``."""
    return _hip.louvain(indptr, indices, data, n, resolution=resolution, random_state=random_state,
                        max_levels=max_levels, commit_prob=commit_prob)


def louvain(connectivities, resolution=1.0, random_state=0, backend="igraph", variant="sync",
            commit_prob=0.9, max_levels=20):
    """Louvain labels (int64 numpy, dense 0..C-1) of a symmetric connectivity matrix.

    ``backend="gpu"`` follows graph.louvain.louvain (src/metalsinglecell/graph/louvain.py:
    397-437) with the synchronous move pass; ``backend="igraph"`` is igraph's
    ``community_multilevel`` (src/metalsinglecell/tl.py:92-102)."""
    if backend == "gpu":
        if variant != "sync":
            raise NotImplementedError(f"variant={variant!r}: only the synchronous variant ('sync') "
                                      "is implemented on the MI355X")
        if not 0.0 < commit_prob <= 1.0:
            raise ValueError("commit_prob must be in (0, 1]")
        if not 1 <= int(max_levels) <= 20:
            raise ValueError("max_levels must be in [1, 20]")
        csr = _canonical_csr(connectivities)
        lab, _, _ = louvain_device(csr.indptr, csr.indices, csr.data, csr.shape[0], resolution,
                                   random_state, max_levels, commit_prob)
        return lab.cpu().numpy().astype(np.int64)
    if backend != "igraph":
        raise ValueError(f"unknown backend {backend!r} (gpu|igraph)")
    try:
        import igraph as ig
    except ImportError as exc:
        raise ImportError("backend='igraph' needs python-igraph, which is not installed; "
                          "use backend='gpu' for the MI355X Louvain") from exc
    import random as _random
    coo = connectivities.tocoo()
    upper = coo.row < coo.col
    try:
        ig.set_random_number_generator(_random.Random(random_state))
        g = ig.Graph(n=connectivities.shape[0],
                     edges=np.column_stack([coo.row[upper], coo.col[upper]]).tolist())
        g.es["weight"] = coo.data[upper].tolist()
        part = g.community_multilevel(weights="weight", resolution=resolution)
    finally:
        ig.set_random_number_generator(_random)
    return np.asarray(part.membership)
