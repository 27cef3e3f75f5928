"""``gr`` namespace — the spatial drop-ins of ``squidpy.gr``: the kNN and fixed-radius spatial
graphs and Moran's I.

Mirrors the reference's ``metalsinglecell.gr`` (src/metalsinglecell/gr.py):

* ``spatial_neighbors_knn`` (:198-225): a directed graph with ``n_neighs`` edges per
  observation, self excluded, neighbours ordered by (distance, index); binary float32
  connectivities with columns in that order, float64 edge lengths from float64
  coordinates; per-library blocks (:107-147); percentile / set_diag / transform tail
  (:37-104); slots ``obsp[f"{key_added}_connectivities"]``, ``obsp[f"{key_added}_distances"]``
  and ``uns[f"{key_added}_neighbors"]`` (:157-168).  The kNN runs in csrc/grid.hip.
* ``spatial_neighbors_radius`` (:228-252): every pair within ``radius`` (or a ``(min, max)``
  interval), self excluded, binary connectivities and float64 lengths, the same tail and
  slots as the kNN graph; pairs from csrc/grid.hip's count / fill passes.
* ``co_occurrence`` (:514-530): the all-pairs cluster co-occurrence histogram in
  csrc/cooccur.hip, squidpy's thresholds and conditional normalisation on the host.
* ``spatial_autocorr`` (:418-503), Moran and Geary modes: statistic and permutation null in
  csrc/moran.hip, the frame ``uns['moranI']`` / ``uns['gearyC']`` with squidpy's columns,
  sorted by I descending / C ascending.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from . import spatial as _sp


class SpatialNeighborsResult(NamedTuple):
    """What the builders return under ``copy=True`` (squidpy's contract)."""

    connectivities: object
    distances: object


def _coords(adata, spatial_key="spatial"):
    if spatial_key not in adata.obsm:
        raise KeyError(f"no adata.obsm[{spatial_key!r}] with spatial coordinates")
    exact = np.ascontiguousarray(adata.obsm[spatial_key], dtype=np.float64)
    return np.ascontiguousarray(exact, dtype=np.float32), exact


def _transform_adj(adj, transform):
    import scipy.sparse as sp
    if transform in (None, "none"):
        return adj
    if transform == "spectral":
        deg = np.asarray(adj.sum(axis=0)).ravel()        # column degree, as squidpy
        inv = np.zeros_like(deg, dtype=np.float64)
        np.divide(1.0, np.sqrt(deg), out=inv, where=deg > 0)
        D = sp.diags(inv)
        return (D @ adj @ D).tocsr()
    if transform == "cosine":
        from sklearn.metrics.pairwise import cosine_similarity
        return sp.csr_matrix(cosine_similarity(adj, dense_output=False))
    raise NotImplementedError(f"transform={transform!r} is not implemented "
                              f"(expected None, 'spectral' or 'cosine')")


def _row_ids(M):
    return np.repeat(np.arange(M.shape[0], dtype=np.int64), np.diff(M.indptr))


def _drop_diag(M):
    import scipy.sparse as sp
    keep = M.indices != _row_ids(M)
    counts = np.bincount(_row_ids(M)[keep], minlength=M.shape[0])
    indptr = np.zeros(M.shape[0] + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    return sp.csr_matrix((M.data[keep], M.indices[keep], indptr), shape=M.shape)


def _finalize(adj, dst, transform, set_diag, percentile):
    """percentile prune on the distances -> set_diag -> drop the distance diagonal -> transform."""
    import scipy.sparse as sp
    adj, dst = adj.tocsr(), dst.tocsr()
    if percentile is not None and dst.nnz:
        keep = dst.data <= np.percentile(dst.data, percentile)
        adj = sp.csr_matrix((adj.data * keep, adj.indices.copy(), adj.indptr.copy()), shape=adj.shape)
        dst = sp.csr_matrix((dst.data * keep, dst.indices.copy(), dst.indptr.copy()), shape=dst.shape)
        adj.eliminate_zeros()
        dst.eliminate_zeros()
    if set_diag:
        adj = adj.maximum(sp.identity(adj.shape[0], format="csr", dtype=adj.dtype)).tocsr()
    dst = _drop_diag(dst)
    dst.eliminate_zeros()
    return _transform_adj(adj, transform), dst


def _library_blocks(adata, library_key):
    import pandas as pd
    if library_key is None:
        return None
    if library_key not in adata.obs:
        raise KeyError(f"no adata.obs[{library_key!r}] to group libraries by")
    col = adata.obs[library_key]
    if not isinstance(col.dtype, pd.CategoricalDtype):
        raise TypeError(f"adata.obs[{library_key!r}] must be categorical (squidpy requires it); "
                        f"got {col.dtype}")
    codes = col.cat.codes.to_numpy()
    blocks = [np.flatnonzero(codes == c) for c in range(len(col.cat.categories))]
    return [b for b in blocks if b.size]


def _assemble(adata, spatial_key, library_key, build, transform, set_diag, percentile):
    import scipy.sparse as sp
    coords, exact = _coords(adata, spatial_key)
    blocks = _library_blocks(adata, library_key)
    if blocks is None:
        return _finalize(*build(coords, exact), transform, set_diag, percentile)
    mats = [_finalize(*build(np.ascontiguousarray(coords[b]), np.ascontiguousarray(exact[b])),
                      transform, set_diag, percentile) for b in blocks]
    adj = sp.block_diag([m[0] for m in mats], format="csr")
    dst = sp.block_diag([m[1] for m in mats], format="csr")
    order = np.concatenate(blocks) if blocks else np.zeros(0, dtype=np.int64)
    if np.any(np.diff(order) < 0):
        inv = np.argsort(order)
        adj, dst = adj[inv][:, inv], dst[inv][:, inv]
    return adj.tocsr(), dst.tocsr()


def _write_graph(adata, adj, dst, params, key_added, transform, copy):
    if copy:
        return SpatialNeighborsResult(adj, dst)
    ck, dk = f"{key_added}_connectivities", f"{key_added}_distances"
    adata.obsp[ck] = adj
    adata.obsp[dk] = dst
    adata.uns[f"{key_added}_neighbors"] = {"connectivities_key": ck, "distances_key": dk,
                                           "params": {**params, "transform": transform}}
    return None


def knn_graph(coords, exact, n_neighs):
    """``(connectivities, distances)`` CSR of the directed spatial kNN graph of one section."""
    import scipy.sparse as sp
    from . import _hip
    n = coords.shape[0]
    k = min(int(n_neighs), n - 1)
    indptr = np.arange(n + 1, dtype=np.int64) * k
    if k <= 0:
        z = sp.csr_matrix((n, n), dtype=np.float32)
        return z, sp.csr_matrix((n, n), dtype=np.float64)
    idx, _, len64 = _hip.spatial_knn(coords, k, coords64=exact)
    cols = idx.cpu().numpy().ravel().astype(np.int64)
    dist = len64.cpu().numpy().ravel()
    return (sp.csr_matrix((np.ones(cols.size, np.float32), cols, indptr), shape=(n, n)),
            sp.csr_matrix((dist, cols, indptr), shape=(n, n)))


def spatial_neighbors_knn(adata, *, n_neighs: int = 6, spatial_key: str = "spatial",
                          library_key: str | None = None, percentile: float | None = None,
                          transform: str | None = None, set_diag: bool = False,
                          key_added: str = "spatial", copy: bool = False):
    """k-NN spatial graph (``sq.gr.spatial_neighbors_knn``) on the MI355X grid index."""
    if transform not in (None, "none", "spectral", "cosine"):
        raise NotImplementedError(f"transform={transform!r} is not implemented "
                                  f"(expected None, 'spectral' or 'cosine')")

    def build(coords, exact):
        return knn_graph(coords, exact, n_neighs)

    adj, dst = _assemble(adata, spatial_key, library_key, build, transform, set_diag, percentile)
    return _write_graph(adata, adj, dst, {"coord_type": "generic", "n_neighbors": n_neighs},
                        key_added, transform, copy)


def radius_graph(coords, exact, r_min, r_max):
    """``(connectivities, distances)`` CSR of the fixed-radius graph of one section."""
    import scipy.sparse as sp
    from . import _hip
    n = coords.shape[0]
    ip, ind, _, len64 = _hip.spatial_radius(coords, r_max, r_min, coords64=exact)
    indptr = ip.cpu().numpy()
    cols = ind.cpu().numpy()
    dist = len64.cpu().numpy()
    return (sp.csr_matrix((np.ones(cols.size, np.float32), cols, indptr), shape=(n, n)),
            sp.csr_matrix((dist, cols.copy(), indptr.copy()), shape=(n, n)))


def spatial_neighbors_radius(adata, *, radius, spatial_key: str = "spatial", library_key: str | None = None,
                             percentile: float | None = None, transform: str | None = None,
                             set_diag: bool = False, key_added: str = "spatial", copy: bool = False):
    """Fixed-radius spatial graph (``sq.gr.spatial_neighbors_radius``) on the MI355X grid.

    ``radius`` is a maximum distance or a ``(min, max)`` interval; every pair within it is
    connected (symmetric graph, degrees follow local density)."""
    if transform not in (None, "none", "spectral", "cosine"):
        raise NotImplementedError(f"transform={transform!r} is not implemented "
                                  f"(expected None, 'spectral' or 'cosine')")
    lo, hi = (0.0, float(radius)) if np.isscalar(radius) else (float(min(radius)), float(max(radius)))

    def build(coords, exact):
        return radius_graph(coords, exact, lo, hi)

    adj, dst = _assemble(adata, spatial_key, library_key, build, transform, set_diag, percentile)
    return _write_graph(adata, adj, dst, {"coord_type": "generic", "radius": radius}, key_added, transform, copy)


def co_occurrence(adata, cluster_key, interval=50, spatial_key: str = "spatial", copy: bool = False):
    """Cluster co-occurrence vs distance (``sq.gr.co_occurrence``, reference gr.py:514-530);
    writes ``uns[f'{cluster_key}_co_occurrence']`` = ``{"occ", "interval"}``.  ``interval`` is
    the number of thresholds or an explicit ascending array of them (``occ`` has one entry
    fewer along its last axis); ``copy=True`` returns ``(occ, interval)``."""
    import pandas as pd
    if cluster_key not in adata.obs:
        raise KeyError(f"no adata.obs[{cluster_key!r}]")
    if not isinstance(adata.obs[cluster_key].dtype, pd.CategoricalDtype):
        raise TypeError(f"expected adata.obs[{cluster_key!r}] to be categorical, got {adata.obs[cluster_key].dtype}")
    if spatial_key not in adata.obsm:
        raise KeyError(f"no adata.obsm[{spatial_key!r}] with spatial coordinates")
    kw = ({"interval": np.asarray(interval, dtype=np.float64)} if np.ndim(interval)
          else {"n_intervals": int(interval)})
    res = _sp.co_occurrence(np.asarray(adata.obsm[spatial_key], dtype=np.float32),
                            adata.obs[cluster_key].to_numpy(), **kw)
    out = {"occ": res["occ"], "interval": res["interval"]}
    if copy:
        return out["occ"], out["interval"]
    adata.uns[f"{cluster_key}_co_occurrence"] = out
    return None


def _autocorr_values(adata, attr, genes, layer, use_raw):
    import scipy.sparse as sp
    if attr == "obs":
        if genes is None:
            df = adata.obs.select_dtypes(include=np.number)
            return np.asarray(df.to_numpy(), dtype=np.float32), df.columns.tolist()
        genes = [genes] if isinstance(genes, str) else list(genes)
        return np.asarray(adata.obs[genes].to_numpy(), dtype=np.float32), genes
    if attr == "obsm":
        if layer is None or layer not in adata.obsm:
            raise KeyError(f"attr='obsm' needs layer= naming a key in adata.obsm, got {layer!r}")
        M = np.asarray(adata.obsm[layer])
        ixs = list(range(M.shape[1])) if genes is None else list(np.ravel([genes]))
        return np.asarray(M[:, ixs], dtype=np.float32), ixs
    if attr != "X":
        raise NotImplementedError(f"extracting from adata.{attr} is not implemented "
                                  f"(expected 'X', 'obs' or 'obsm')")
    if use_raw and getattr(adata, "raw", None) is None:
        raise AttributeError("no .raw attribute found; try use_raw=False")
    source = adata.raw if use_raw else adata
    if genes is None:
        genes = (adata.var_names[np.asarray(adata.var["highly_variable"], dtype=bool)]
                 if "highly_variable" in adata.var else adata.var_names).tolist()
    genes = [genes] if isinstance(genes, str) else list(genes)
    if use_raw:
        keep = set(source.var_names)
        genes = [g for g in genes if g in keep]
    gi = [source.var_names.get_loc(g) for g in genes]
    X = source.X if (use_raw or layer is None) else adata.layers[layer]
    Xg = X[:, gi]
    return np.asarray(Xg.todense() if sp.issparse(Xg) else Xg, dtype=np.float32), genes


def moran_frame(score, sims, g, index, two_tailed=False, corr_method="fdr_bh", mode="moran"):
    """The ``uns['moranI']`` / ``uns['gearyC']`` frame from the statistic, the permutation null
    and the weight matrix (gr.py:484-498)."""
    import pandas as pd
    stat = "I" if mode == "moran" else "C"
    p_norm, var_norm = _sp.analytic_pval(score, g, mode=mode, two_tailed=two_tailed)
    cols = {stat: score, "pval_norm": p_norm, "var_norm": np.full(score.shape, var_norm)}
    if sims is not None:
        p_sim, p_z_sim, var_sim = _sp.permutation_pvals(score, sims)
        cols.update(pval_z_sim=p_z_sim, pval_sim=p_sim, var_sim=var_sim)
    df = pd.DataFrame(cols, index=index)
    if corr_method is not None:
        for pv in [c for c in df.columns if "pval" in c]:
            df[f"{pv}_{corr_method}"] = _sp.multiple_testing(df[pv].to_numpy(), corr_method)
    df.sort_values(by=stat, ascending=(mode == "geary"), inplace=True)
    return df


def spatial_autocorr(adata, mode: str = "moran", genes=None, n_perms: int | None = None,
                     connectivity_key: str = "spatial_connectivities", layer=None,
                     transformation: bool = True, two_tailed: bool = False,
                     corr_method: str | None = "fdr_bh", attr: str = "X", use_raw: bool = False,
                     seed: int = 0, copy: bool = False, permutations: str = "numpy"):
    """Moran's I / Geary's C per gene (``sq.gr.spatial_autocorr``); writes ``uns['moranI']`` /
    ``uns['gearyC']``.

    ``permutations="numpy"`` (default) draws the null from ``default_rng(seed)`` as the
    reference does; ``"device"`` generates it on the GPU (no host permutation stream).
    """
    if mode not in ("moran", "geary"):
        raise ValueError(f"mode must be 'moran' or 'geary', got {mode!r}")
    Xg, index = _autocorr_values(adata, attr, genes, layer, use_raw)
    g = adata.obsp[connectivity_key]
    g = _sp.row_normalize(g) if transformation else g.astype(np.float64)
    out = _sp.spatial_autocorr(Xg, g, mode=mode, n_perms=n_perms, random_state=seed,
                               permutations=permutations)
    df = moran_frame(out[mode], out["sims"], g, index, two_tailed, corr_method, mode=mode)
    if copy:
        return df
    adata.uns["moranI" if mode == "moran" else "gearyC"] = df
    return None
