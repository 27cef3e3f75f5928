"""``gr`` namespace — the spatial drop-ins of ``squidpy.gr``: the kNN and fixed-radius spatial
graphs, co-occurrence and spatial autocorrelation (Moran's I, Geary's C).

Public surface and slots follow the reference's ``metalsinglecell.gr`` (src/metalsinglecell/
gr.py): ``spatial_neighbors_knn`` (:198-225), ``spatial_neighbors_radius`` (:228-252),
``co_occurrence`` (:514-530) and ``spatial_autocorr`` (:455-503).  The neighbours come from
csrc/grid.hip, the statistics from csrc/moran.hip and csrc/cooccur.hip.

Graph contract (squidpy's): a cell's out-edges go to its ``n_neighs`` nearest other cells by
(float32 distance, index), or to every cell within the radius; connectivities are 1.0
(float32), distances the float64 edge lengths of the float64 coordinates.  With
``library_key`` each section (category of that column) is a graph of its own, and the tail —
``percentile`` (keep the edges whose length is at most that percentile of the section's edge
lengths), ``set_diag`` (a 1.0 self-edge in the connectivities) and ``transform`` — acts per
section.  The distance matrix never stores the diagonal or zero lengths.

How it is built here: every section's directed edges go into one global edge list (cell ids
of the whole object), pruned by the section's own percentile; the two CSR matrices are then
formed once from that list, rows in cell order and each row's edges in the order the device
emitted them (distance order for kNN, column order for the radius graph), a self-edge first.
The transform runs once on the assembled matrix: both transforms keep the block structure
(the spectral scale uses column degrees, which no edge crosses between sections; a cosine
between rows of different sections has no common column), so one call equals the per-section
calls entry for entry.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

from . import spatial as _sp

_TRANSFORMS = (None, "none", "spectral", "cosine")


class SpatialNeighborsResult(NamedTuple):
    """What the builders return under ``copy=True`` (squidpy's contract)."""

    connectivities: object
    distances: object


# ------------------------------------------------------------------------ graph assembly
def _require_transform(transform):
    if transform not in _TRANSFORMS:
        raise NotImplementedError(f"transform={transform!r} is not implemented "
                                  f"(expected None, 'spectral' or 'cosine')")


def _section_members(adata, library_key, n):
    """Cell ids of each section in category order (empty categories skipped); ``[None]``
    when the whole object is one section."""
    import pandas as pd
    if library_key is None:
        return [None]
    if library_key not in adata.obs:
        raise KeyError(f"no adata.obs[{library_key!r}] to group libraries by")
    column = adata.obs[library_key]
    if not isinstance(column.dtype, pd.CategoricalDtype):
        raise TypeError(f"adata.obs[{library_key!r}] must be categorical (squidpy requires it); "
                        f"got {column.dtype}")
    code = np.asarray(column.cat.codes, dtype=np.int64)
    if (code < 0).any():
        raise ValueError(f"adata.obs[{library_key!r}] has cells without a library")
    by_code = np.argsort(code, kind="stable")
    bounds = np.searchsorted(code[by_code], np.arange(len(column.cat.categories) + 1))
    return [by_code[a:b] for a, b in zip(bounds[:-1], bounds[1:]) if b > a]


def _edges_to_csr(n, src, dst, values, self_value=None):
    """CSR from an edge list; a row keeps its edges in list order.  ``self_value`` puts one
    self-edge of that value at the front of every row."""
    import scipy.sparse as sp
    if self_value is not None:
        every = np.arange(n, dtype=np.int64)
        src = np.concatenate([every, src])
        dst = np.concatenate([every, dst])
        values = np.concatenate([np.full(n, self_value, dtype=values.dtype), values])
    order = np.argsort(src, kind="stable")
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=indptr[1:])
    return sp.csr_matrix((values[order], dst[order], indptr), shape=(n, n))


def _spectral(adj):
    """``D^-1/2 A D^-1/2`` with D the COLUMN degrees (squidpy sums axis 0); a cell that no
    edge reaches scales by 0 (squidpy divides by zero there), and zeros are not stored."""
    import scipy.sparse as sp
    m = sp.csr_matrix(adj, dtype=np.float64)
    col_deg = np.bincount(m.indices, weights=m.data, minlength=m.shape[1])
    scale = np.zeros(m.shape[1])
    np.divide(1.0, np.sqrt(col_deg), out=scale, where=col_deg > 0)
    rows = np.repeat(np.arange(m.shape[0]), np.diff(m.indptr))
    out = sp.csr_matrix(((m.data * scale[rows]) * scale[m.indices], m.indices, m.indptr),
                        shape=m.shape)
    out.eliminate_zeros()
    return out


def _apply_transform(adj, transform):
    if transform in (None, "none"):
        return adj
    if transform == "spectral":
        return _spectral(adj)
    import scipy.sparse as sp
    from sklearn.metrics.pairwise import cosine_similarity
    return sp.csr_matrix(cosine_similarity(adj, dense_output=False))


def _section_graphs(adata, spatial_key, library_key, edges_of, transform, set_diag, percentile):
    """``(connectivities, distances)`` of the whole object from the per-section builder
    ``edges_of(xy32, xy64) -> (indptr, local column ids, float64 lengths)``."""
    _require_transform(transform)
    if spatial_key not in adata.obsm:
        raise KeyError(f"no adata.obsm[{spatial_key!r}] with spatial coordinates")
    xy64 = np.ascontiguousarray(adata.obsm[spatial_key], dtype=np.float64)
    n = xy64.shape[0]
    src, dst, length = [], [], []
    for ids in _section_members(adata, library_key, n):
        sub64 = xy64 if ids is None else np.ascontiguousarray(xy64[ids])
        indptr, cols, lens = edges_of(sub64.astype(np.float32), sub64)
        rows = np.repeat(np.arange(sub64.shape[0], dtype=np.int64), np.diff(indptr))
        cols = np.asarray(cols, dtype=np.int64)
        if percentile is not None and lens.size:
            cut = lens <= np.percentile(lens, percentile)
            rows, cols, lens = rows[cut], cols[cut], lens[cut]
        if ids is not None:
            rows, cols = ids[rows], ids[cols]
        src.append(rows)
        dst.append(cols)
        length.append(np.asarray(lens, dtype=np.float64))
    src, dst, length = np.concatenate(src), np.concatenate(dst), np.concatenate(length)
    adj = _edges_to_csr(n, src, dst, np.ones(src.size, np.float32),
                        self_value=np.float32(1.0) if set_diag else None)
    proper = (src != dst) & (length != 0.0)
    dist = _edges_to_csr(n, src[proper], dst[proper], length[proper])
    return _apply_transform(adj, transform), dist


def _store_graph(adata, adj, dist, key_added, params, copy):
    if copy:
        return SpatialNeighborsResult(adj, dist)
    slots = {"connectivities_key": f"{key_added}_connectivities",
             "distances_key": f"{key_added}_distances"}
    adata.obsp[slots["connectivities_key"]] = adj
    adata.obsp[slots["distances_key"]] = dist
    adata.uns[f"{key_added}_neighbors"] = dict(slots, params=params)
    return None


# ------------------------------------------------------------------------ builders
def knn_edges(xy32, xy64, n_neighs):
    """Directed kNN edges of one section on the device grid: ``(indptr, cols, float64
    lengths)``, each row's neighbours in (float32 distance, index) order.  A section of
    m <= n_neighs cells links every cell to the other m - 1."""
    from . import _hip
    m = xy32.shape[0]
    k = min(int(n_neighs), m - 1)
    if k <= 0:
        return np.zeros(m + 1, np.int64), np.zeros(0, np.int64), np.zeros(0)
    idx, _, len64 = _hip.spatial_knn(xy32, k, coords64=xy64)
    return (np.arange(m + 1, dtype=np.int64) * k, idx.cpu().numpy().ravel(),
            len64.cpu().numpy().ravel())


def radius_edges(xy32, xy64, r_min, r_max):
    """Every pair of one section with r_min <= distance <= r_max, self excluded:
    ``(indptr, cols, float64 lengths)``, columns ascending."""
    from . import _hip
    ip, ind, _, len64 = _hip.spatial_radius(xy32, r_max, r_min, coords64=xy64)
    return ip.cpu().numpy(), ind.cpu().numpy(), len64.cpu().numpy()


def spatial_neighbors_knn(adata, *, n_neighs: int = 6, spatial_key: str = "spatial",
                          library_key: str | None = None, percentile: float | None = None,
                          transform: str | None = None, set_diag: bool = False,
                          key_added: str = "spatial", copy: bool = False):
    """k-NN spatial graph (``sq.gr.spatial_neighbors_knn``) on the MI355X grid index."""
    adj, dist = _section_graphs(adata, spatial_key, library_key,
                                lambda a, b: knn_edges(a, b, n_neighs), transform, set_diag,
                                percentile)
    params = {"coord_type": "generic", "n_neighbors": n_neighs, "transform": transform}
    return _store_graph(adata, adj, dist, key_added, params, copy)


def spatial_neighbors_radius(adata, *, radius, spatial_key: str = "spatial", library_key: str | None = None,
                             percentile: float | None = None, transform: str | None = None,
                             set_diag: bool = False, key_added: str = "spatial", copy: bool = False):
    """Fixed-radius spatial graph (``sq.gr.spatial_neighbors_radius``) on the MI355X grid.

    ``radius`` is a maximum distance or a ``(min, max)`` interval; every pair within it is
    connected (symmetric graph, degrees follow local density)."""
    if np.isscalar(radius):
        r_min, r_max = 0.0, float(radius)
    else:
        r_min, r_max = float(np.min(radius)), float(np.max(radius))
    adj, dist = _section_graphs(adata, spatial_key, library_key,
                                lambda a, b: radius_edges(a, b, r_min, r_max), transform,
                                set_diag, percentile)
    params = {"coord_type": "generic", "radius": radius, "transform": transform}
    return _store_graph(adata, adj, dist, key_added, params, copy)


# ------------------------------------------------------------------------ co-occurrence
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


# ------------------------------------------------------------------------ autocorrelation
def _dense32(block):
    import scipy.sparse as sp
    return np.asarray(block.toarray() if sp.issparse(block) else block, dtype=np.float32)


def _values_from_obs(adata, genes, layer, use_raw):
    """Numeric ``obs`` columns: all of them, or the named ones."""
    frame = adata.obs.select_dtypes(include=np.number) if genes is None else \
        adata.obs[[genes] if isinstance(genes, str) else list(genes)]
    return _dense32(frame.to_numpy()), list(frame.columns)


def _values_from_obsm(adata, genes, layer, use_raw):
    """Columns of ``obsm[layer]``: all of them, or the given positions."""
    if layer is None or layer not in adata.obsm:
        raise KeyError(f"attr='obsm' needs layer= naming a key in adata.obsm, got {layer!r}")
    block = np.asarray(adata.obsm[layer])
    cols = np.arange(block.shape[1]) if genes is None else np.atleast_1d(np.asarray(genes)).ravel()
    return _dense32(block[:, cols]), cols.tolist()


def _values_from_var(adata, genes, layer, use_raw):
    """Genes of ``.X`` (or ``layers[layer]``, or ``.raw``): the named ones, else the highly
    variable ones when ``var['highly_variable']`` exists, else all.  With ``use_raw`` the
    names missing from ``.raw`` are dropped, keeping the order."""
    if use_raw and getattr(adata, "raw", None) is None:
        raise AttributeError("no .raw attribute found; try use_raw=False")
    if genes is None:
        names = adata.var_names
        if "highly_variable" in adata.var:
            names = names[np.asarray(adata.var["highly_variable"], dtype=bool)]
        wanted = list(names)
    else:
        wanted = [genes] if isinstance(genes, str) else list(genes)
    holder = adata.raw if use_raw else adata
    position = {name: p for p, name in enumerate(holder.var_names)}
    if use_raw:
        wanted = [g for g in wanted if g in position]
    matrix = holder.X if (use_raw or layer is None) else adata.layers[layer]
    return _dense32(matrix[:, [position[g] for g in wanted]]), wanted


_VALUE_SOURCES = {"X": _values_from_var, "obs": _values_from_obs, "obsm": _values_from_obsm}


def moran_frame(score, sims, g, index, two_tailed=False, corr_method="fdr_bh", mode="moran"):
    """The ``uns['moranI']`` / ``uns['gearyC']`` frame (squidpy's layout): the statistic,
    ``pval_norm``, ``var_norm``, with a permutation null also ``pval_z_sim``, ``pval_sim``,
    ``var_sim``; then one ``<p-column>_<corr_method>`` per p-value column in that order.
    Rows are sorted by the statistic (I descending, C ascending; NaN last, ties in input
    order)."""
    import pandas as pd
    stat = "I" if mode == "moran" else "C"
    p_norm, var_norm = _sp.analytic_pval(score, g, mode=mode, two_tailed=two_tailed)
    columns = [(stat, score), ("pval_norm", p_norm), ("var_norm", np.full(score.shape, var_norm))]
    if sims is not None:
        p_sim, p_z_sim, var_sim = _sp.permutation_pvals(score, sims)
        columns += [("pval_z_sim", p_z_sim), ("pval_sim", p_sim), ("var_sim", var_sim)]
    if corr_method is not None:
        columns += [(f"{name}_{corr_method}", _sp.multiple_testing(np.asarray(v), corr_method))
                    for name, v in list(columns) if name.startswith("pval")]
    key = np.asarray(score, dtype=np.float64)
    key = np.where(np.isnan(key), np.inf, key if mode == "geary" else -key)
    order = np.argsort(key, kind="stable")
    return pd.DataFrame({name: np.asarray(v)[order] for name, v in columns},
                        index=pd.Index(list(index)).take(order))


def spatial_autocorr(adata, mode: str = "moran", genes=None, n_perms: int | None = None,
                     connectivity_key: str = "spatial_connectivities", layer=None,
                     transformation: bool = True, two_tailed: bool = False,
                     corr_method: str | None = "fdr_bh", attr: str = "X", use_raw: bool = False,
                     seed: int = 0, copy: bool = False, permutations: str = "numpy"):
    """Moran's I / Geary's C per gene (``sq.gr.spatial_autocorr``); writes ``uns['moranI']`` /
    ``uns['gearyC']`` (``copy=True`` returns the frame instead).

    ``permutations="numpy"`` (default) draws the null from ``default_rng(seed)`` as the
    reference does; ``"device"`` generates it on the GPU (no host permutation stream).
    """
    if mode not in ("moran", "geary"):
        raise ValueError(f"mode must be 'moran' or 'geary', got {mode!r}")
    if attr not in _VALUE_SOURCES:
        raise NotImplementedError(f"extracting from adata.{attr} is not implemented "
                                  f"(expected 'X', 'obs' or 'obsm')")
    values, names = _VALUE_SOURCES[attr](adata, genes, layer, use_raw)
    weights = adata.obsp[connectivity_key]
    weights = _sp.row_normalize(weights) if transformation else weights.astype(np.float64)
    result = _sp.spatial_autocorr(values, weights, mode=mode, n_perms=n_perms, random_state=seed,
                                  permutations=permutations)
    frame = moran_frame(result[mode], result["sims"], weights, names, two_tailed, corr_method,
                        mode=mode)
    if copy:
        return frame
    adata.uns["moranI" if mode == "moran" else "gearyC"] = frame
    return None
