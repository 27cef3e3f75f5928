"""Synchronous parallel Leiden oracle — TEST INFRASTRUCTURE ONLY.

Restates the reference's ``variant="sync"`` Leiden (src/metalsinglecell/graph/leiden.py:
251-455 with graph/louvain.py:88-131, 282-394, 440-453, graph/primitives.py:94-111,
graph/csr_graph.py:24-63), vectorised in NumPy.  Per pass:

1. ``sigtot[c] = sum_{v: comm[v]=c} k_v`` (louvain.py:357).
2. Best move of every active vertex (louvain.py:88-131): over neighbours u != v,
   ``w_{v->c}`` = summed edge weight into community c, ``sig' = sigtot[c] - [c==curr] k_v``,
   ``gain = w - gamma * sig' * k_v / 2m``; stay gain = gain(curr) if a neighbour is in curr,
   else 0; best gain wins, ties to the smaller c; move only if best > stay + 1e-9.
   Refinement (leiden.py:100-144) restricts neighbours to ``part[u] == part[v]`` and starts
   from singletons.
3. ``wants = target != comm``; ``coin = u32 < floor(0.9 * 2^32)``; commit ``wants & coin``.
4. Next active set = commit | N(commit) | (wants & ~coin) (louvain.py:364-372), on every
   graph (see the last deviation below).
5. Convergence test every 4 passes (and at the last): stop when nothing wants to move.

Levels (leiden.py:410-452): move from P; relabel densely (np.unique order); reject the level
when Q <= Q_prev + 1e-9; stop if every node is its own community; refine; stop if refinement
is the identity; aggregate on the refined partition (intra edges become self-loops, weights
summed) with P lifted; at most 20 levels.  ``n_iterations`` repeats the pass with
``seed + 17 * it`` from the previous labels (leiden.py:374-377).

``louvain`` restates the multilevel Louvain of the same module (louvain.py:397-453) on the
same move pass and contraction: levels move from singletons, no refinement.

Deliberate definitions shared bit-for-bit with the HIP kernels (documented deviations from
the reference's MLX arithmetic):
* edge weights are quantised once to 64-bit fixed point, ``wq = rint(w * 2^s)`` with
  ``s = 52 - ceil(log2(max w)) - ceil(log2(E))`` (both exact, order-independent), so every
  degree, sigma_tot and w_{v->c} is an exact integer sum below 2^52; gains are then
  evaluated in float64 by a fixed expression and the modularity sum runs sequentially over
  communities in label order;
* the coin comes from the counter RNG in oracle/numerics.py keyed by (seed, phase, pass, v)
  instead of MLX's key/split stream;
* vertices of degree > 1024 follow the same synchronous rule as all others (the reference
  moves them sequentially on the host, louvain.py:151-181); since no vertex is moved outside
  the synchronous pass, the frontier rule applies to every graph (the reference switches
  pruning off when such vertices exist because its host moves bypass the frontier,
  louvain.py:350-352).
"""

from __future__ import annotations

import numpy as np

from .numerics import coin_threshold, coin_u32, stream_key

MAX_LEVELS = 20
TOL = 1e-9


class QGraph:
    """CSR graph with fixed-point weights (sorted columns, both directions stored)."""

    def __init__(self, indptr, indices, wq, scale_exp):
        self.indptr = np.asarray(indptr, np.int64)
        self.indices = np.asarray(indices, np.int64)
        self.wq = np.asarray(wq, np.int64)
        self.s = int(scale_exp)
        self.n = self.indptr.size - 1
        self.src = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.indptr))
        self.k = np.bincount(self.src, weights=self.wq.astype(np.float64),
                             minlength=self.n).astype(np.int64)
        self.twom = int(self.wq.sum())
        self.max_deg = int(np.diff(self.indptr).max()) if self.n else 0

    @property
    def inv_scale(self):
        return float(np.ldexp(1.0, -self.s))


def ceil_log2(x):
    """Exact ceil(log2(x)) for x > 0 (from the binary exponent)."""
    import math
    m, e = math.frexp(float(x))
    return e - 1 if m == 0.5 else e


def weight_scale_exponent(max_w, n_edges):
    """``s`` with ``max_w * n_edges * 2^s <= 2^52`` (0 for an empty or all-zero graph)."""
    if n_edges <= 0 or max_w <= 0:
        return 0
    return int(52 - ceil_log2(max_w) - ceil_log2(n_edges))


def quantize(w, s):
    return np.rint(np.asarray(w, np.float64) * np.ldexp(1.0, s)).astype(np.int64)


def from_scipy(mat):
    """``Graph.from_scipy`` (csr_graph.py:24-39): canonical CSR, float32 weights, quantised."""
    import scipy.sparse as sp
    csr = sp.csr_matrix(mat)
    csr.sum_duplicates()
    csr.sort_indices()
    w = csr.data.astype(np.float32)
    s = weight_scale_exponent(float(w.max()) if w.size else 0.0, w.size)
    return QGraph(csr.indptr, csr.indices, quantize(w, s), s)


def _segsum_sorted(keys_sorted, vals_sorted):
    starts = np.flatnonzero(np.r_[True, keys_sorted[1:] != keys_sorted[:-1]])
    return keys_sorted[starts], np.add.reduceat(vals_sorted, starts), starts


def best_moves(g, comm, sigtot, active, gamma, part=None):
    """Target community per vertex (the move / refine kernel, louvain.py:88-131)."""
    n = g.n
    target = comm.copy()
    src, dst, w = g.src, g.indices, g.wq
    ok = (dst != src) & active[src]
    if part is not None:
        ok &= part[dst] == part[src]
    es, ed, ew = src[ok], dst[ok], w[ok]
    if es.size == 0:
        return target
    key = es * n + comm[ed]
    order = np.argsort(key, kind="stable")
    ukey, wc, _ = _segsum_sorted(key[order], ew[order])
    v = ukey // n
    c = ukey % n
    curr = comm[v]
    inv_s = g.inv_scale
    inv = 1.0 / (float(g.twom) * inv_s)
    sig = sigtot[c] - np.where(c == curr, g.k[v], 0)
    gain = wc.astype(np.float64) * inv_s - ((gamma * (sig.astype(np.float64) * inv_s))
                                          * (g.k[v].astype(np.float64) * inv_s)) * inv
    vstart = np.flatnonzero(np.r_[True, v[1:] != v[:-1]])
    best = np.maximum.reduceat(gain, vstart)
    seg = np.repeat(np.arange(vstart.size), np.diff(np.r_[vstart, v.size]))
    is_best = gain == best[seg]
    # first position of the max in each segment (segments are sorted by c ascending)
    cand = np.where(is_best, np.arange(v.size), v.size)
    first = np.minimum.reduceat(cand, vstart)
    best_c = c[first]
    stay_mask = c == curr
    stay = np.zeros(vstart.size)
    stay[seg[stay_mask]] = gain[stay_mask]
    vv = v[vstart]
    move = best > stay + TOL
    target[vv[move]] = best_c[move]
    return target


def sigma_tot(g, comm):
    return np.bincount(comm, weights=g.k.astype(np.float64), minlength=g.n).astype(np.int64)


def local_moving_sync(g, gamma, seed, init_comm, part=None, max_passes=200,
                      commit_prob=0.9, sync_every=4):
    """``_local_moving_sync`` (louvain.py:282-394) / ``_refine_sync`` (leiden.py:251-349)."""
    n = g.n
    comm = np.asarray(init_comm, np.int64).copy()
    key = stream_key(seed, 0 if part is None else 1)
    thr = coin_threshold(commit_prob)
    active = np.ones(n, bool)
    vid = np.arange(n, dtype=np.uint64)
    passes = 0
    for p in range(max_passes):
        passes = p + 1
        sig = sigma_tot(g, comm)
        target = best_moves(g, comm, sig, active, gamma, part)
        wants = target != comm
        coin = coin_u32(key, p, vid) < thr
        commit = wants & coin
        comm = np.where(commit, target, comm)
        nbr = np.zeros(n, bool)
        if g.src.size:
            hit = commit[g.indices]
            nbr = np.bincount(g.src[hit], minlength=n) > 0
        active = commit | nbr | (wants & ~coin)
        if (p % sync_every == sync_every - 1) or (p == max_passes - 1):
            if not wants.any():
                break
    return comm, passes


def dense_labels(comm):
    _, inv = np.unique(comm, return_inverse=True)
    return inv.astype(np.int64).ravel()


def modularity(g, comm, gamma):
    """Q = sum_c [in_c / 2m - gamma (tot_c / 2m)^2] (primitives.py:94-111), float64 sum."""
    C = int(comm.max()) + 1 if comm.size else 0
    tot = np.bincount(comm, weights=g.k.astype(np.float64), minlength=C)
    same = comm[g.src] == comm[g.indices]
    inw = np.bincount(comm[g.src][same], weights=g.wq[same].astype(np.float64), minlength=C)
    twom = float(g.twom) * g.inv_scale
    tot = tot * g.inv_scale
    inw = inw * g.inv_scale
    a = tot / twom
    terms = inw / twom - gamma * (a * a)
    q = 0.0
    for t in terms.tolist():          # sequential, label order (the GPU sums the same way)
        q += t
    return q


def contract(g, R, nR):
    """``_contract_dense`` (louvain.py:440-453): exact integer sums of (R[src], R[dst])."""
    key = R[g.src] * nR + R[g.indices]
    order = np.argsort(key, kind="stable")
    ukey, w, _ = _segsum_sorted(key[order], g.wq[order])
    s_new = ukey // nR
    d_new = ukey % nR
    indptr = np.zeros(nR + 1, np.int64)
    np.cumsum(np.bincount(s_new, minlength=nR), out=indptr[1:])
    return QGraph(indptr, d_new, w, g.s)


def leiden_pass(g0, gamma, seed, init, commit_prob=0.9, max_levels=MAX_LEVELS, stats=None):
    """``_leiden_pass`` (leiden.py:381-455)."""
    g = g0
    orig2node = np.arange(g0.n, dtype=np.int64)
    P = np.asarray(init, np.int64).copy() if init is not None else np.arange(g0.n, dtype=np.int64)
    q_prev = -1.0
    final = P[orig2node]
    for level in range(max_levels):
        prev_labels = P[orig2node]
        comm, mp = local_moving_sync(g, gamma, seed + level, P, commit_prob=commit_prob)
        P = dense_labels(comm)
        n_comms = int(P.max()) + 1
        q = modularity(g, P, gamma)
        if stats is not None:
            stats.append(dict(level=level, n=g.n, E=g.indices.size, move_passes=mp,
                              n_comms=n_comms, q=q))
        if q <= q_prev + TOL:
            final = prev_labels
            break
        q_prev = q
        final = P[orig2node]
        if n_comms == g.n:
            break
        Rc, rp = local_moving_sync(g, gamma, seed + level, np.arange(g.n), part=P,
                                   commit_prob=commit_prob)
        R = dense_labels(Rc)
        nR = int(R.max()) + 1
        if stats is not None:
            stats[-1].update(refine_passes=rp, n_refined=nR)
        if nR == g.n:
            break
        P_agg = np.zeros(nR, np.int64)
        P_agg[R] = P
        orig2node = R[orig2node]
        g = contract(g, R, nR)
        P = P_agg
    return dense_labels(final)


def leiden(conn, resolution=1.0, random_state=0, n_iterations=2, commit_prob=0.9, stats=None):
    """``graph.leiden.leiden`` (leiden.py:352-378) on a symmetric scipy connectivity matrix."""
    g = from_scipy(conn)
    labels = None
    for it in range(max(1, n_iterations)):
        labels = leiden_pass(g, resolution, random_state + 17 * it, labels,
                             commit_prob=commit_prob, stats=stats)
    return labels


def modularity_of_labels(conn, labels, resolution=1.0):
    """Modularity of a labelling on the quantised graph (float64)."""
    g = from_scipy(conn)
    return modularity(g, dense_labels(np.asarray(labels)), resolution)


def louvain(conn, resolution=1.0, random_state=0, max_levels=MAX_LEVELS, commit_prob=0.9, stats=None):
    """Multilevel Louvain, ``graph.louvain.louvain`` with ``variant="sync"`` (louvain.py:397-437).

    Level l: synchronous local moving from singletons with seed ``random_state + 100 l``;
    dense relabel; stop, keeping the previous level's labels, when every vertex stayed alone
    or Q <= Q_prev + 1e-9; otherwise map the original vertices through the moved partition
    and aggregate on it (``_contract_dense``, louvain.py:440-453).  Returns dense labels."""
    g = from_scipy(conn)
    orig2super = np.arange(g.n, dtype=np.int64)
    q_prev = -1.0
    for level in range(max_levels):
        comm, mp = local_moving_sync(g, resolution, random_state + 100 * level, np.arange(g.n),
                                     commit_prob=commit_prob)
        cd = dense_labels(comm)
        C = int(cd.max()) + 1
        q = modularity(g, cd, resolution)
        if stats is not None:
            stats.append(dict(level=level, n=g.n, E=g.indices.size, move_passes=mp, n_comms=C, q=q))
        if C == g.n or q <= q_prev + TOL:
            break
        orig2super = cd[orig2super]
        g = contract(g, cd, C)
        q_prev = q
    return dense_labels(orig2super)
