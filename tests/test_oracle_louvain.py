"""Pins the Louvain oracle (oracle/leiden.py::louvain): a closed-form optimum, the two
resolution limits, brute-force optimal modularity on tiny graphs, an independent optimiser
(networkx), and the multilevel invariants."""

import networkx as nx
import numpy as np
import scipy.sparse as sp
from sklearn.metrics import adjusted_rand_score

from oracle import leiden


def _csr(G):
    return sp.csr_matrix(nx.to_scipy_sparse_array(G, nodelist=range(G.number_of_nodes()),
                                                  weight="weight", format="csr").astype(np.float32))


def test_ring_of_cliques_closed_form():
    """m cliques of size s on a ring: the clique partition has Q = m (l/L - (d/2L)^2) with
    l = s(s-1)/2 intra edges, L = m l + m edges and d = 2 l + 2 degree per clique."""
    m, s = 12, 6
    A = _csr(nx.ring_of_cliques(m, s))
    out = leiden.louvain(A)
    assert adjusted_rand_score(np.arange(m * s) // s, out) == 1.0
    l = s * (s - 1) / 2
    L = m * l + m
    q_closed = m * (l / L - ((2 * l + 2) / (2 * L)) ** 2)
    assert abs(leiden.modularity_of_labels(A, out) - q_closed) < 1e-12


def _components_graph(seed=0):
    rng = np.random.default_rng(seed)
    blocks = []
    for b, size in enumerate([7, 12, 20, 5]):
        s = b
        H = nx.gnp_random_graph(size, 0.5, seed=s)
        while not nx.is_connected(H):
            s += 100
            H = nx.gnp_random_graph(size, 0.5, seed=s)
        blocks.append(H)
    U = nx.disjoint_union_all(blocks)
    for u, v in U.edges:
        U[u][v]["weight"] = float(rng.uniform(0.2, 1.0))
    return _csr(U)


def test_resolution_limits():
    """gamma = 0: every move into a neighbour community gains, so the levels merge each
    connected component into one community (Q = 1).  gamma huge: no move gains, so every
    vertex stays alone and the labels are the identity."""
    A = _components_graph()
    cc = sp.csgraph.connected_components(A, directed=False)[1]
    out = leiden.louvain(A, resolution=0.0)
    assert adjusted_rand_score(cc, out) == 1.0
    assert abs(leiden.modularity_of_labels(A, out, 0.0) - 1.0) < 1e-12
    assert np.array_equal(leiden.louvain(A, resolution=1e6), np.arange(A.shape[0]))


def _set_partitions(n):
    a = [0] * n

    def rec(i, m):
        if i == n:
            yield list(a)
            return
        for c in range(m + 1):
            a[i] = c
            yield from rec(i + 1, m + 1 if c == m else m)

    yield from rec(0, 0)


def test_brute_force_optimum_tiny_graphs():
    """On two planted groups of 9 vertices Louvain reaches the optimum over all 21147 set
    partitions; on every graph its Q never exceeds that optimum."""
    for seed in range(3):
        r = np.random.default_rng(seed)
        n = 9
        W = np.zeros((n, n), np.float32)
        for i in range(n):
            for j in range(i + 1, n):
                if r.uniform() < (0.8 if (i < 4) == (j < 4) else 0.15):
                    W[i, j] = W[j, i] = r.uniform(0.1, 1.0)
        A = sp.csr_matrix(W)
        g = leiden.from_scipy(A)
        best = max(leiden.modularity(g, leiden.dense_labels(np.array(p)), 1.0) for p in _set_partitions(n))
        q = leiden.modularity_of_labels(A, leiden.louvain(A))
        assert q <= best + 1e-12
        assert abs(q - best) < 1e-12


def test_planted_partition_vs_networkx_louvain():
    sizes = [80, 100, 120, 60]
    G = nx.stochastic_block_model(sizes, [[0.2 if i == j else 0.01 for j in range(4)] for i in range(4)], seed=3)
    rng = np.random.default_rng(3)
    for u, v in G.edges:
        G[u][v]["weight"] = float(rng.uniform(0.2, 1.0))
    A = _csr(G)
    out = leiden.louvain(A)
    assert adjusted_rand_score(np.repeat(np.arange(4), sizes), out) > 0.95
    q_nx = nx.community.modularity(G, nx.community.louvain_communities(G, weight="weight", seed=0),
                                   weight="weight")
    assert leiden.modularity_of_labels(A, out) >= q_nx - 0.02


def test_levels_strictly_improve_and_contraction_preserves_q():
    """Accepted levels raise Q by more than 1e-9 each; the reported Q of the last accepted
    level is the Q of the returned labels; and the singleton partition of the contracted
    graph has exactly the Q of the partition it was contracted on (exact integer weights)."""
    G = nx.stochastic_block_model([40] * 10, [[0.2 if i == j else 0.01 for j in range(10)] for i in range(10)],
                                  seed=7)
    A = _csr(G)
    stats = []
    out = leiden.louvain(A, stats=stats)
    assert set(np.unique(out)) == set(range(out.max() + 1))
    # the last record is the level that stopped the run (nothing moved or no gain)
    *accepted, last = stats
    assert accepted and (last["n_comms"] == last["n"] or last["q"] <= accepted[-1]["q"] + 1e-9)
    for a, b in zip(accepted, accepted[1:]):
        assert b["q"] > a["q"] + 1e-9 and b["n_comms"] < b["n"]
    assert abs(leiden.modularity_of_labels(A, out) - accepted[-1]["q"]) < 1e-12
    g = leiden.from_scipy(A)
    R = leiden.dense_labels(out)
    gc = leiden.contract(g, R, int(R.max()) + 1)
    assert leiden.modularity(gc, np.arange(gc.n), 1.0) == leiden.modularity(g, R, 1.0)


def test_max_levels_one_is_the_first_move_phase():
    """max_levels=1 returns the first level's moved partition when it improves on -1."""
    A = _components_graph(1)
    out1 = leiden.louvain(A, max_levels=1)
    comm, _ = leiden.local_moving_sync(leiden.from_scipy(A), 1.0, 0, np.arange(A.shape[0]))
    assert np.array_equal(out1, leiden.dense_labels(comm))
