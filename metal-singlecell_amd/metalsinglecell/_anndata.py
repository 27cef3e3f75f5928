"""A minimal AnnData-compatible container.

The ``pp``/``tl``/``gr`` functions are duck-typed: they use ``obs``, ``var``, ``obsm``,
``obsp``, ``uns``, ``layers``, ``X``, ``n_obs``, ``n_vars``, ``var_names`` and ``copy()``,
so a real ``anndata.AnnData`` works unchanged.  This class provides exactly that surface
for environments where anndata is not installed (tests, benchmarks, serving jobs).
"""

from __future__ import annotations

import copy as _copy

import numpy as np


class AnnData:
    def __init__(self, X=None, obs=None, var=None, obsm=None, obsp=None, uns=None, layers=None,
                 n_obs=None):
        import pandas as pd
        self.X = X
        n = X.shape[0] if X is not None else int(n_obs or 0)
        g = X.shape[1] if X is not None else 0
        self.obs = obs if obs is not None else pd.DataFrame(index=[str(i) for i in range(n)])
        self.var = var if var is not None else pd.DataFrame(index=[f"g{i}" for i in range(g)])
        self.obsm = dict(obsm or {})
        self.obsp = dict(obsp or {})
        self.uns = dict(uns or {})
        self.layers = dict(layers or {})
        self.raw = None

    @property
    def n_obs(self):
        return len(self.obs.index)

    @property
    def n_vars(self):
        return len(self.var.index)

    @property
    def shape(self):
        return (self.n_obs, self.n_vars)

    @property
    def var_names(self):
        return self.var.index

    @var_names.setter
    def var_names(self, names):
        self.var.index = [str(x) for x in names]

    @property
    def obs_names(self):
        return self.obs.index

    def copy(self):
        return _copy.deepcopy(self)

    def __getitem__(self, mask):
        """Row subset (boolean mask or integer index) of X, obs, obsm and obsp."""
        idx = np.flatnonzero(mask) if np.asarray(mask).dtype == bool else np.asarray(mask)
        X = None if self.X is None else self.X[idx]
        out = AnnData(X=X, obs=self.obs.iloc[idx].copy(), var=self.var.copy(),
                      uns=_copy.deepcopy(self.uns))
        out.obsm = {k: np.asarray(v)[idx] for k, v in self.obsm.items()}
        out.obsp = {k: v[idx][:, idx] for k, v in self.obsp.items()}
        out.layers = {k: v[idx] for k, v in self.layers.items()}
        return out

    def __repr__(self):
        return (f"AnnData(n_obs={self.n_obs}, n_vars={self.n_vars}; obsm={list(self.obsm)}, "
                f"obsp={list(self.obsp)}, uns={list(self.uns)})")
