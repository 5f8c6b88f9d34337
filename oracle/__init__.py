"""CPU oracle for the group Steiner tree solver -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() (as the checker) and bench.py's CPU-baseline leg use this
package.  It wraps oracle/liboracle.so (dpbf_oracle.c, built by oracle/Makefile) and adds the
closed forms that pin it:

* g = 1: cost 0;  g = 2: shortest group-to-group path = min_c d(c,G1) + d(c,G2);
* g = 3: a minimal tree touching 3 groups is a path or a spider, so
  OPT = min_c sum_i d(c, G_i);
* g = 4: the minimal subtree spanning one representative per group has at most two branch
  vertices, so OPT = min over the 3 pairings {A,B} of min_{c1,c2} S_A(c1) + d(c1,c2) + S_B(c2)
  with S_A(c) = sum_{i in A} d(c, G_i); the inner minimum is one Dijkstra seeded with S_A.

Functions take any object with ``row_ptr``/``col``/``w`` numpy arrays (a symmetric CSR) and groups
with ``g``/``members(i)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P = C.c_void_p
    lib.orc_dpbf.restype = C.c_int
    lib.orc_dpbf.argtypes = [C.c_int32, P, P, P, C.c_int32, P, P, P, P, P, P, C.c_int32, P, P]
    lib.orc_dijkstra.restype = C.c_int
    lib.orc_dijkstra.argtypes = [C.c_int32, P, P, P, P, P]
    lib.orc_bruteforce.restype = C.c_double
    lib.orc_bruteforce.argtypes = [C.c_int32, P, P, P, C.c_int32, P]
    lib.orc_dpbf_pruned.restype = C.c_int
    lib.orc_dpbf_pruned.argtypes = [C.c_int32, P, P, P, C.c_int32, P, C.c_double, C.c_double, C.c_int32, P, P, P]
    lib.orc_dpbf_batch.restype = C.c_int
    lib.orc_state_bounds.restype = C.c_int
    lib.orc_state_bounds.argtypes = [C.c_int32, C.c_int32, P, P, C.c_int32, P]
    lib.orc_dpbf_batch.argtypes = [C.c_int32, C.c_int32, P, P, P, C.c_int32, P, P, P]
    _lib = lib
    return lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else 0


def _csr(graph):
    row = np.ascontiguousarray(graph.row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(graph.col, dtype=np.int32)
    w = np.ascontiguousarray(graph.w, dtype=np.float64)
    return row, col, w


def group_masks(n: int, groups) -> np.ndarray:
    gm = np.zeros(n, dtype=np.uint32)
    for i in range(groups.g):
        gm[np.asarray(groups.members(i))] |= np.uint32(1 << i)
    return gm


def dpbf(graph, groups):
    """Label-setting DPBF.  Returns (cost, root, eu, ev, ew, stats); cost = inf if infeasible."""
    lib = load()
    row, col, w = _csr(graph)
    n = row.shape[0] - 1
    gm = group_masks(n, groups)
    cap = n + 64
    eu = np.zeros(cap, np.int32)
    ev = np.zeros(cap, np.int32)
    ew = np.zeros(cap, np.float64)
    cost = C.c_double(0)
    root = C.c_int32(-1)
    ne = C.c_int32(0)
    stats = np.zeros(3, np.int64)
    rc = lib.orc_dpbf(n, _p(row), _p(col), _p(w), groups.g, _p(gm), C.addressof(cost), C.addressof(root),
                      _p(eu), _p(ev), _p(ew), cap, C.addressof(ne), _p(stats))
    if rc not in (0, 1):
        raise RuntimeError(f"orc_dpbf failed with code {rc}")
    k = min(ne.value, cap)
    return cost.value, root.value, eu[:k].copy(), ev[:k].copy(), ew[:k].copy(), \
        {"pops": int(stats[0]), "relax": int(stats[1]), "merges": int(stats[2])}


def dpbf_cost(graph, groups) -> float:
    return dpbf(graph, groups)[0]


def dpbf_pruned(graph, groups, ub: float, grow_frac: float = 0.5, use_lb=True, pop_ub: bool = False, dist=None):
    """Label-setting DPBF under a feasible cost ub (PrunedDP++ bounds): only states of cost
    <= grow_frac * ub are grown (0: all); use_lb: False, True / 1 (one-label), 2 (one-label and
    tour bounds), 3 (tour bounds with a planted mistake, for the tests); pop_ub: PrunedDP++'s
    feasible tree at every pop (popped tree + shortest paths to the missing groups) lowers ub, and
    the search ends once the popped cost reaches it.  dist: group_distances(graph, groups) when the
    caller has them.  Returns (cost, stats)."""
    lib = load()
    row, col, w = _csr(graph)
    n = row.shape[0] - 1
    gm = group_masks(n, groups)
    use_lb = int(use_lb)
    if use_lb or pop_ub:
        dist = np.ascontiguousarray(group_distances(graph, groups) if dist is None else dist, dtype=np.float64)
    else:
        dist = np.zeros(1)
    if pop_ub:
        use_lb |= 16
    cost = C.c_double(0)
    stats = np.zeros(3, np.int64)
    rc = lib.orc_dpbf_pruned(n, _p(row), _p(col), _p(w), groups.g, _p(gm), float(ub), float(grow_frac),
                             use_lb, _p(dist), C.addressof(cost), _p(stats))
    if rc not in (0, 1):
        raise RuntimeError(f"orc_dpbf_pruned failed with code {rc}")
    return cost.value, {"pops": int(stats[0]), "relax": int(stats[1]), "merges": int(stats[2])}


def state_bounds(graph, groups, use_lb: int = 2) -> np.ndarray:
    """lb[v, p]: the lower bound that dpbf_pruned(use_lb=...) applies to state (v, p)."""
    lib = load()
    n = graph.row_ptr.shape[0] - 1
    gm = group_masks(n, groups)
    dist = np.ascontiguousarray(group_distances(graph, groups))
    out = np.empty((n, 1 << groups.g), np.float64)
    rc = lib.orc_state_bounds(n, groups.g, _p(gm), _p(dist), int(use_lb), _p(out))
    if rc != 0:
        raise RuntimeError(f"orc_state_bounds failed with code {rc}")
    return out


def dijkstra(graph, init: np.ndarray) -> np.ndarray:
    lib = load()
    row, col, w = _csr(graph)
    n = row.shape[0] - 1
    init = np.ascontiguousarray(init, dtype=np.float64)
    dist = np.empty(n, np.float64)
    lib.orc_dijkstra(n, _p(row), _p(col), _p(w), _p(init), _p(dist))
    return dist


def group_distances(graph, groups) -> np.ndarray:
    """d[i, v] = distance from v to the nearest member of group i."""
    n = graph.row_ptr.shape[0] - 1
    out = np.empty((groups.g, n), np.float64)
    for i in range(groups.g):
        init = np.full(n, np.inf)
        init[np.asarray(groups.members(i))] = 0.0
        out[i] = dijkstra(graph, init)
    return out


def closed_form(graph, groups) -> float:
    """Optimal cost for g <= 4 from shortest paths alone (see module docstring)."""
    g = groups.g
    d = group_distances(graph, groups)
    if g == 1:
        return 0.0 if len(groups.members(0)) else np.inf
    if g in (2, 3):
        return float(np.min(d.sum(axis=0)))
    if g == 4:
        best = np.inf
        for A, B in (((0, 1), (2, 3)), ((0, 2), (1, 3)), ((0, 3), (1, 2))):
            sa = d[A[0]] + d[A[1]]
            sb = d[B[0]] + d[B[1]]
            da = dijkstra(graph, sa)
            best = min(best, float(np.min(da + sb)))
        return best
    raise ValueError("closed form only for g <= 4")


def bruteforce(graph, groups) -> float:
    lib = load()
    row, col, w = _csr(graph)
    n = row.shape[0] - 1
    if n > 20:
        raise ValueError("brute force needs n <= 20")
    gm = group_masks(n, groups)
    return float(lib.orc_bruteforce(n, _p(row), _p(col), _p(w), groups.g, _p(gm)))


def dpbf_batch(graph, group_sets, nthreads: int = 1) -> np.ndarray:
    """orc_dpbf over many group sets on the same graph with a thread pool; returns costs."""
    lib = load()
    row, col, w = _csr(graph)
    n = row.shape[0] - 1
    nq = len(group_sets)
    gq = np.array([gs.g for gs in group_sets], dtype=np.int32)
    gms = np.stack([group_masks(n, gs) for gs in group_sets]).astype(np.uint32)
    costs = np.empty(nq, np.float64)
    lib.orc_dpbf_batch(nthreads, n, _p(row), _p(col), _p(w), nq, _p(gq), _p(np.ascontiguousarray(gms)), _p(costs))
    return costs


def mst_induced(graph, vertices):
    """Minimum spanning forest of the subgraph induced by `vertices`: restates the MST post-processing
    rucgraph/graph_hash_of_mixed_weighted_MST_postprocessing.h:73-103 (the induced subgraph of the
    vertex set, then the Prim of graph_hash_of_mixed_weighted/common_algorithms/
    graph_hash_of_mixed_weighted_minimum_spanning_tree.h:29-132, here restarted in every component, so
    that a disconnected set gives a spanning forest).  Returns (edges as (u, v, w) with u < v, cost)."""
    import heapq
    vs = sorted(set(int(x) for x in vertices))
    inset = set(vs)
    rp, col, w = graph.row_ptr, graph.col, graph.w
    adj = {v: [] for v in vs}
    for v in vs:
        best = {}
        for e in range(int(rp[v]), int(rp[v + 1])):
            x = int(col[e])
            if x in inset and x != v:
                wx = float(w[e])
                if x not in best or wx < best[x]:
                    best[x] = wx
        adj[v] = sorted(best.items())
    done, edges, cost = set(), [], 0.0
    for s in vs:
        if s in done:
            continue
        done.add(s)
        heap = [(wx, x, s) for x, wx in adj[s]]
        heapq.heapify(heap)
        while heap:
            wx, x, p = heapq.heappop(heap)
            if x in done:
                continue
            done.add(x)
            edges.append((min(x, p), max(x, p), wx))
            cost += wx
            for y, wy in adj[x]:
                if y not in done:
                    heapq.heappush(heap, (wy, y, x))
    return edges, cost


def spider_cost(graph, groups) -> float:
    """The automatic bound before post-processing: min over centres c of sum_i d(c, G_i)."""
    d = group_distances(graph, groups)
    return float(np.min(d.sum(axis=0)))
