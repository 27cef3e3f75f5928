"""metalsinglecell for MI355X — the neighbor-graph path of metal-SingleCell on gfx950.

``import metalsinglecell as msc`` gives the reference's API for this path:

* ``msc.pp.neighbors``            exact kNN + UMAP connectivities
* ``msc.pp.bbknn``                batch-balanced exact kNN + connectivities
* ``msc.tl.leiden(backend="gpu")`` synchronous parallel Leiden
* ``msc.tl.louvain(backend="gpu")`` synchronous parallel multilevel Louvain
* ``msc.gr.spatial_neighbors_knn`` exact grid spatial kNN
* ``msc.gr.spatial_autocorr``      Moran's I with a permutation null

The compute runs in ``_lib/libmsc_hip.so`` (hand-written HIP kernels for CDNA4).  Importing
the package loads neither torch nor the library; both load on the first GPU call.
"""

from . import gr, pp, tl  # noqa: F401

__version__ = "0.1.0"
__all__ = ["pp", "tl", "gr"]
