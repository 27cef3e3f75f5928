"""``tl`` namespace — the Leiden and Louvain drop-ins of ``scanpy.tl.leiden`` / ``tl.louvain``.

Mirrors the reference's ``metalsinglecell.tl.leiden`` (src/metalsinglecell/tl.py:55-73) and
``tl.louvain`` (:76-105) with
``_conn`` (:14-21) and ``_categorical`` (:35-52): the connectivity graph is resolved through
``uns[neighbors_key]``, labels are written to ``obs[key_added]`` as a Categorical of string
labels with natural-sorted categories, and ``uns[key_added]['params']`` records resolution,
random_state and n_iterations (Louvain: resolution and random_state).  ``backend="gpu"`` runs
csrc/leiden.hip on the MI355X.
"""

from __future__ import annotations

import re

import numpy as np


def _conn(adata, neighbors_key="neighbors"):
    key = "connectivities"
    if neighbors_key in adata.uns:
        key = adata.uns[neighbors_key].get("connectivities_key", key)
    if key not in adata.obsp:
        raise ValueError(f"run msc.pp.neighbors first (no adata.obsp[{key!r}])")
    return adata.obsp[key]


def _natkey(s):
    """Natural-order key: digit runs compare as integers."""
    return [int(t) if t.isdigit() else t for t in re.split(r"(\d+)", str(s))]


def _categorical(labels):
    """Categorical of the labels' strings with natural-sorted categories, built from the distinct
    labels (one string each) and integer codes, not from a string per cell."""
    import pandas as pd
    uniq, inv = np.unique(np.asarray(labels), return_inverse=True)
    names = [str(x) for x in uniq.tolist()]
    cats = sorted(set(names), key=_natkey)
    where = {c: i for i, c in enumerate(cats)}
    code_of_uniq = np.fromiter((where[nm] for nm in names), dtype=np.int64, count=len(names))
    return pd.Categorical.from_codes(code_of_uniq[np.asarray(inv).reshape(-1)], categories=cats)


def leiden(adata, resolution: float = 1.0, key_added: str = "leiden", random_state: int = 0,
           n_iterations: int = 2, backend: str = "igraph", variant: str = "sync",
           commit_prob: float = 0.9, neighbors_key: str = "neighbors", copy: bool = False):
    """Leiden clustering; writes ``obs[key_added]`` (categorical) and ``uns[key_added]``."""
    from .cluster import leiden as _leiden
    adata = adata.copy() if copy else adata
    lab = _leiden(_conn(adata, neighbors_key), resolution=resolution, random_state=random_state,
                  n_iterations=n_iterations, backend=backend, variant=variant,
                  commit_prob=commit_prob)
    adata.obs[key_added] = _categorical(lab)
    adata.uns[key_added] = {"params": {"resolution": resolution, "random_state": random_state,
                                       "n_iterations": n_iterations}}
    return adata if copy else None


def louvain(adata, resolution: float = 1.0, key_added: str = "louvain", random_state: int = 0,
            backend: str = "igraph", variant: str = "sync", commit_prob: float = 0.9,
            neighbors_key: str = "neighbors", copy: bool = False):
    """Louvain clustering; writes ``obs[key_added]`` (categorical) and ``uns[key_added]``."""
    from .cluster import louvain as _louvain
    adata = adata.copy() if copy else adata
    lab = _louvain(_conn(adata, neighbors_key), resolution=resolution, random_state=random_state,
                   backend=backend, variant=variant, commit_prob=commit_prob)
    adata.obs[key_added] = _categorical(lab)
    adata.uns[key_added] = {"params": {"resolution": resolution, "random_state": random_state}}
    return adata if copy else None
