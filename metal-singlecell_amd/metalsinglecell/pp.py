"""``pp`` namespace — the neighbour-graph drop-ins of ``scanpy.pp.neighbors`` and
``sc.external.pp.bbknn``.

Mirrors the reference's ``metalsinglecell.pp.neighbors`` (src/metalsinglecell/pp.py:753-788)
and its representation choice (:722-745): same signature (``n_pcs`` positional after
``n_neighbors``, ``use_rep`` keyword-only), same slots (``obsp[distances_key]``,
``obsp[connectivities_key]``, ``uns[key]`` with ``params``), same ``copy`` semantics.
The kNN is exact at every size and runs on the MI355X (csrc/knn.hip, csrc/fuzzy.hip).
The front-end PCA the reference computes when ``X_pca`` is missing on a wide object is
outside this build: that case raises and asks for ``obsm['X_pca']``.
"""

from __future__ import annotations

import numpy as np

_N_PCS_DEFAULT = 50


def _dense(X):
    import scipy.sparse as sp
    return X.toarray() if sp.issparse(X) else np.asarray(X)


def _choose_representation(adata, use_rep, n_pcs):
    """scanpy's representation rule (reference pp.py:722-745).

    ``use_rep=None``: a narrow object (at most 50 variables) is clustered on ``.X`` whether or
    not it holds ``X_pca``; a wide one on ``obsm['X_pca']``, which must have at least
    ``n_pcs or 50`` columns.  Where the reference would compute that PCA itself, this build
    raises: the PCA front-end is not part of the neighbour-graph path.
    """
    wide = adata.n_vars > _N_PCS_DEFAULT
    if use_rep is None and (n_pcs == 0 or not wide):
        return np.asarray(_dense(adata.X), dtype=np.float32), None
    if use_rep is None:
        need = n_pcs or _N_PCS_DEFAULT
        have = np.asarray(adata.obsm["X_pca"]).shape[1] if "X_pca" in adata.obsm else 0
        if have < need:
            raise NotImplementedError(
                f"{adata.n_vars} variables and obsm['X_pca'] has {have} < {need} columns: "
                "compute the PCA first (the PCA front-end is not part of this MI355X "
                "neighbour-graph build)")
        return np.asarray(adata.obsm["X_pca"], dtype=np.float32), "X_pca"
    if use_rep == "X":
        return np.asarray(_dense(adata.X), dtype=np.float32), "X"
    if use_rep not in adata.obsm:
        raise ValueError(f"Did not find {use_rep} in .obsm.keys(). You need to compute it first.")
    return np.asarray(adata.obsm[use_rep], dtype=np.float32), use_rep


def neighbors(adata, n_neighbors: int = 15, n_pcs: int | None = None, *, use_rep: str | None = None,
              metric: str = "euclidean", key_added: str | None = None, random_state: int = 0,
              copy: bool = False, precision: str = "bf16"):
    """kNN graph; writes ``obsp['distances']``, ``obsp['connectivities']``, ``uns['neighbors']``.

    ``precision`` selects the MFMA form that ranks candidates (``"bf16"`` or ``"fp32"``);
    both re-rank exactly and return identical graphs.
    """
    from .neighbors import neighbors as _nb
    if metric != "euclidean":
        raise NotImplementedError(f"metric={metric!r} is not implemented (Euclidean only)")
    adata = adata.copy() if copy else adata
    rep, _ = _choose_representation(adata, use_rep, n_pcs)
    if n_pcs is not None:
        rep = rep[:, :n_pcs]
    dist, conn = _nb(np.ascontiguousarray(rep), n_neighbors=n_neighbors,
                     random_state=random_state, precision=precision)
    ck = "connectivities" if key_added is None else f"{key_added}_connectivities"
    dk = "distances" if key_added is None else f"{key_added}_distances"
    adata.obsp[dk] = dist
    adata.obsp[ck] = conn
    params = {"n_neighbors": n_neighbors, "method": "umap", "random_state": random_state,
              "metric": metric}
    if use_rep is not None:
        params["use_rep"] = use_rep
    if n_pcs is not None:
        params["n_pcs"] = n_pcs
    adata.uns["neighbors" if key_added is None else key_added] = {
        "connectivities_key": ck, "distances_key": dk, "params": params}
    return adata if copy else None


def bbknn(adata, batch_key: str = "batch", *, use_rep: str = "X_pca", neighbors_within_batch: int = 3,
          n_pcs: int | None = 50, trim: int | None = None, random_state: int = 0, copy: bool = False,
          precision: str = "bf16"):
    """Batch-balanced kNN (``sc.external.pp.bbknn``, reference pp.py:809-833); writes
    ``obsp['distances']``, ``obsp['connectivities']`` and ``uns['neighbors']`` whose
    ``n_neighbors`` is the combined width K = sum over batches of min(k, batch size)."""
    from .neighbors import bbknn as _bbknn
    adata = adata.copy() if copy else adata
    if use_rep not in adata.obsm:
        raise ValueError(f"Did not find {use_rep} in .obsm.keys(). You need to compute it first.")
    if batch_key not in adata.obs:
        raise KeyError(f"batch_key {batch_key!r} not in adata.obs")
    dist, conn, K = _bbknn(np.asarray(adata.obsm[use_rep], dtype=np.float32), np.asarray(adata.obs[batch_key]),
                           neighbors_within_batch=neighbors_within_batch, n_pcs=n_pcs, trim=trim,
                           random_state=random_state, precision=precision)
    adata.obsp["distances"] = dist
    adata.obsp["connectivities"] = conn
    adata.uns["neighbors"] = {
        "connectivities_key": "connectivities", "distances_key": "distances",
        "params": {"n_neighbors": K, "method": "umap", "metric": "euclidean",
                   "random_state": random_state, "use_rep": use_rep, "n_pcs": n_pcs,
                   "bbknn": {"trim": 10 * K if trim is None else trim, "computation": "hip"}}}
    return adata if copy else None
