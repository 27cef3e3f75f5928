"""CPU oracle for the neighbor-graph path — TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy/SciPy (and one C file for the brute-force kNN), the
reference algorithm of each hot-path function, citing the reference file:line it follows.
It is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.  The product (``metalsinglecell``) never imports,
links or runs anything from here; its GPU path fails loudly when the HIP library is
missing instead of falling back to this code.

Modules
-------
``numerics``   the project-defined float32 ``expf`` and the counter-based RNG shared by the
               oracle and the HIP kernels (defined once, restated on both sides)
``knn``        exact Euclidean kNN (neighbors.py:420-462 ordering rule, self pinned first)
``fuzzy``      UMAP fuzzy-simplicial-set connectivities (umap-learn 0.5.x, neighbors.py:742-753)
``leiden``     the synchronous parallel Leiden (graph/leiden.py:251-455, graph/louvain.py:88-453)
``grid``       exact 2-D/3-D spatial kNN (neighbors.py:333-417, gr.py:171-225)
``moran``      Moran's I and its permutation / analytic statistics (spatial.py:268-406)
"""
