"""Regenerate the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

Each fixture holds small seeded inputs and the oracle's outputs for one stage of the
neighbor-graph path.  tests/test_golden.py checks the oracle (CPU) and the HIP path (GPU)
against them.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "metal-singlecell_amd")]

from metalsinglecell import synthetic  # noqa: E402
from oracle import fuzzy, grid, knn, leiden, moran  # noqa: E402


def lattice(side, jitter=0.0, seed=0):
    r, c = np.divmod(np.arange(side * side), side)
    xy = np.c_[c * 100.0 + (r % 2) * 50.0, r * 86.6]
    if jitter:
        xy = xy + np.random.default_rng(seed).normal(0, jitter, xy.shape)
    return xy, r, c


def main():
    X, _ = synthetic.gaussian_mixture_isotropic(n=600, d=50, n_clusters=6, seed=10)
    idx, d2, dist = knn.exact_knn(X, 15)
    np.savez_compressed(os.path.join(HERE, "knn_mixture600_k15.npz"), X=X,
                        idx=idx.astype(np.int16), d2=d2)
    P, sigma, rho = fuzzy.fuzzy_simplicial_set(idx, dist)
    np.savez_compressed(os.path.join(HERE, "fuzzy_mixture600_k15.npz"), indptr=P.indptr,
                        indices=P.indices.astype(np.int16), data=P.data, sigma=sigma, rho=rho)
    lab = leiden.leiden(P, resolution=1.0, random_state=0, n_iterations=2)
    q = leiden.modularity_of_labels(P, lab)
    np.savez_compressed(os.path.join(HERE, "leiden_mixture600.npz"), labels=lab.astype(np.int8),
                        q=np.float64(q))
    lab = leiden.louvain(P, resolution=1.0, random_state=0)
    q = leiden.modularity_of_labels(P, lab)
    np.savez_compressed(os.path.join(HERE, "louvain_mixture600.npz"), labels=lab.astype(np.int8),
                        q=np.float64(q))
    xy, r, c = lattice(30, jitter=1.0, seed=0)
    si, sd = grid.spatial_knn(xy.astype(np.float32), 6)
    np.savez_compressed(os.path.join(HERE, "spatial_hex30_k6.npz"), xy=xy,
                        idx=si.astype(np.int16), dist=sd)
    xy, r, c = lattice(25)
    rng = np.random.default_rng(0)
    G = np.c_[(r + c).astype(np.float32), ((r + c) % 2).astype(np.float32),
              rng.normal(size=625)].astype(np.float32)
    A, _ = grid.knn_graph(xy, 6)
    perms = moran.permutations(625, 99, 0)
    I, sims = moran.moran_dense(G, moran.row_normalize(A), perms)
    np.savez_compressed(os.path.join(HERE, "moran_hex25.npz"), xy=xy, X=G, I=I, sims=sims)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
