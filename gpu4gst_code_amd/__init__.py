"""MI355X group Steiner tree (DPBF) solver: Python binding of the C-ABI in include/gst.h.

The solver itself is native: libgst.so (host C++ + gfx950 HIP kernels).  This module only moves
numpy arrays across the ABI for tests, benches and scripts.  Every GPU entry point raises when
the library or the device is missing; there is no CPU fallback.

    import gpu4gst_code_amd as gst
    G = gst.gen_barabasi_albert(200_000, 5, 1, 1000, seed=1)
    groups = gst.gen_groups(G.n, 6, 10, 200, seed=1000)
    with gst.Solver() as s:
        dg = s.upload(G)
        r = s.solve(dg, groups)
        print(r.cost, r.edges)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from ._abi import (GST_MAX_GROUPS, GST_W_FLOAT32, GST_W_INT32, STATUS_NAMES)

__all__ = ["HostGraph", "Groups", "Result", "Solver", "GstError", "gen_random_connected",
           "gen_barabasi_albert", "gen_rmat", "gen_groups", "groups_from_lists", "validate_tree",
           "GST_MAX_GROUPS", "GST_W_INT32", "GST_W_FLOAT32"]


class GstError(RuntimeError):
    def __init__(self, status: int, message: str = ""):
        self.status = status
        self.name = STATUS_NAMES.get(status, f"status {status}")
        super().__init__(f"{self.name}: {message}" if message else self.name)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else 0


@dataclass
class HostGraph:
    """Symmetric CSR on the host (row_ptr int64, col int32, w int32 or float32)."""
    row_ptr: np.ndarray
    col: np.ndarray
    w: np.ndarray

    def __post_init__(self):
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int64)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        if self.w.dtype == np.float32:
            self.w = np.ascontiguousarray(self.w, dtype=np.float32)
        else:
            self.w = np.ascontiguousarray(self.w, dtype=np.int32)

    @property
    def n(self) -> int:
        return int(self.row_ptr.shape[0] - 1)

    @property
    def m(self) -> int:
        return int(self.row_ptr[-1])

    @property
    def w_type(self) -> int:
        return GST_W_FLOAT32 if self.w.dtype == np.float32 else GST_W_INT32

    def as_c(self) -> _abi.gst_csr:
        return _abi.gst_csr(self.n, self.m, _ptr(self.row_ptr), _ptr(self.col), _ptr(self.w), self.w_type)

    @staticmethod
    def from_edges(n: int, eu, ev, w) -> "HostGraph":
        """Undirected edge list -> symmetric CSR (rows sorted by neighbour id)."""
        eu = np.asarray(eu, dtype=np.int64)
        ev = np.asarray(ev, dtype=np.int64)
        w = np.asarray(w)
        src = np.concatenate([eu, ev])
        dst = np.concatenate([ev, eu])
        ww = np.concatenate([w, w])
        order = np.lexsort((dst, src))
        src, dst, ww = src[order], dst[order], ww[order]
        row = np.zeros(n + 1, dtype=np.int64)
        np.add.at(row, src + 1, 1)
        row = np.cumsum(row)
        return HostGraph(row, dst.astype(np.int32), ww.astype(np.float32 if w.dtype.kind == "f" else np.int32))


@dataclass
class Groups:
    """g groups as a CSR of member lists; group i is bit i of the DP masks."""
    grp_ptr: np.ndarray
    grp_vtx: np.ndarray

    def __post_init__(self):
        self.grp_ptr = np.ascontiguousarray(self.grp_ptr, dtype=np.int32)
        self.grp_vtx = np.ascontiguousarray(self.grp_vtx, dtype=np.int32)

    @property
    def g(self) -> int:
        return int(self.grp_ptr.shape[0] - 1)

    def members(self, i: int) -> np.ndarray:
        return self.grp_vtx[self.grp_ptr[i]:self.grp_ptr[i + 1]]

    def as_c(self) -> _abi.gst_groups:
        return _abi.gst_groups(self.g, _ptr(self.grp_ptr), _ptr(self.grp_vtx))

    def masks(self, n: int) -> np.ndarray:
        """Per-vertex group bitmask (uint32)."""
        gm = np.zeros(n, dtype=np.uint32)
        for i in range(self.g):
            gm[self.members(i)] |= np.uint32(1 << i)
        return gm


def groups_from_lists(lists: Sequence[Sequence[int]]) -> Groups:
    ptr = np.zeros(len(lists) + 1, dtype=np.int32)
    for i, l in enumerate(lists):
        ptr[i + 1] = ptr[i] + len(l)
    vtx = np.concatenate([np.asarray(l, dtype=np.int32) for l in lists]) if lists else np.zeros(0, np.int32)
    return Groups(ptr, vtx)


@dataclass
class Result:
    cost: float
    tree_cost: float
    root: int
    eu: np.ndarray
    ev: np.ndarray
    ew: np.ndarray
    used_ub: bool
    fell_back: bool
    stats: dict = field(default_factory=dict)

    @property
    def edges(self):
        return list(zip(self.eu.tolist(), self.ev.tolist(), self.ew.tolist()))


def _check(lib, status: int, ctx=None, what: str = ""):
    if status != _abi.GST_OK:
        msg = lib.gst_last_error(ctx).decode() if ctx else what
        raise GstError(status, msg)


def _take_host_csr(lib, h: _abi.gst_host_csr) -> HostGraph:
    n, m = h.n, h.m
    row = np.ctypeslib.as_array(h.row_ptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(h.col, shape=(max(m, 1),))[:m].copy()
    if h.w_type == GST_W_FLOAT32:
        w = np.ctypeslib.as_array(C.cast(h.w, C.POINTER(C.c_float)), shape=(max(m, 1),))[:m].copy()
    else:
        w = np.ctypeslib.as_array(C.cast(h.w, C.POINTER(C.c_int32)), shape=(max(m, 1),))[:m].copy()
    lib.gst_host_csr_free(C.byref(h))
    return HostGraph(row, col, w)


def gen_random_connected(n: int, n_edges: int, wlo: int, whi: int, seed: int) -> HostGraph:
    """Random spanning tree plus uniform extra edges (no duplicates, no self-loops), int weights."""
    lib = _abi.load()
    h = _abi.gst_host_csr()
    _check(lib, lib.gst_gen_random_connected(n, n_edges, wlo, whi, seed & (2**64 - 1), C.byref(h)),
           what="gen_random_connected")
    return _take_host_csr(lib, h)


def gen_barabasi_albert(n: int, m: int, wlo: int, whi: int, seed: int) -> HostGraph:
    lib = _abi.load()
    h = _abi.gst_host_csr()
    _check(lib, lib.gst_gen_barabasi_albert(n, m, wlo, whi, seed & (2**64 - 1), C.byref(h)),
           what="gen_barabasi_albert")
    return _take_host_csr(lib, h)


def gen_rmat(scale: int, edge_factor: int, a=0.57, b=0.19, c=0.19, *, weights="int", wlo=1, whi=1000,
             mu=0.0, sigma=1.0, seed: int = 0) -> HostGraph:
    lib = _abi.load()
    h = _abi.gst_host_csr()
    wt = GST_W_FLOAT32 if weights == "lognormal" else GST_W_INT32
    _check(lib, lib.gst_gen_rmat(scale, edge_factor, a, b, c, wt, wlo, whi, mu, sigma, seed & (2**64 - 1),
                                 C.byref(h)), what="gen_rmat")
    return _take_host_csr(lib, h)


def gen_groups(n: int, g: int, smin: int, smax: int, seed: int) -> Groups:
    lib = _abi.load()
    ptr = np.zeros(g + 1, dtype=np.int32)
    vtx = np.zeros(g * min(smax, n), dtype=np.int32)
    _check(lib, lib.gst_gen_groups(n, g, smin, smax, seed & (2**64 - 1), _ptr(ptr), _ptr(vtx)), what="gen_groups")
    return Groups(ptr, vtx[:ptr[-1]].copy())


def validate_tree(graph: HostGraph, groups: Groups, eu, ev, root: int = -1) -> float:
    """Raise GstError unless (eu, ev) is a tree of the graph touching every group; return its cost."""
    lib = _abi.load()
    eu = np.ascontiguousarray(eu, dtype=np.int32)
    ev = np.ascontiguousarray(ev, dtype=np.int32)
    cost = C.c_double(0.0)
    csr, grp = graph.as_c(), groups.as_c()
    st = lib.gst_validate_tree(C.byref(csr), C.byref(grp), int(eu.shape[0]), _ptr(eu), _ptr(ev), int(root),
                               C.byref(cost))
    _check(lib, st, what="validate_tree")
    return cost.value


class DeviceGraph:
    def __init__(self, solver: "Solver", handle: int, host: HostGraph):
        self.solver = solver
        self.handle = handle
        self.host = host

    @property
    def n(self) -> int:
        return self.host.n

    def free(self):
        if self.handle:
            self.solver.lib.gst_graph_free(self.handle)
            self.handle = None


class Solver:
    """One gst_ctx on one device (default device 0)."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        ctx = C.c_void_p()
        st = self.lib.gst_ctx_create(device, C.byref(ctx))
        if st != _abi.GST_OK:
            msg = self.lib.gst_last_error(ctx).decode() if ctx.value else ""
            if ctx.value:
                self.lib.gst_ctx_destroy(ctx)
            raise GstError(st, msg)
        self.ctx = ctx
        self._graphs = []

    def close(self):
        for g in self._graphs:
            g.free()
        self._graphs = []
        if self.ctx:
            self.lib.gst_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, graph: HostGraph) -> DeviceGraph:
        h = C.c_void_p()
        csr = graph.as_c()
        _check(self.lib, self.lib.gst_graph_upload(self.ctx, C.byref(csr), C.byref(h)), self.ctx)
        dg = DeviceGraph(self, h.value, graph)
        self._graphs.append(dg)
        return dg

    def solve(self, graph: DeviceGraph, groups: Groups, *, ub: Optional[float] = None, prune_lb: bool = True,
              want_tree: bool = True, rounds_per_sync: int = 0, chunk_arcs: int = 0,
              pull_permille: int = 0, grow_mode: int = 0, push_lanes: int = 0, half_bound: bool = True,
              auto_bound: bool = False, merge_tile_kb: int = 0, tour_bound: bool = True) -> Result:
        opts = _abi.gst_opts()
        opts.ub = float(ub) if ub is not None else 0.0
        opts.prune_lb = 1 if prune_lb else 0
        opts.want_tree = 1 if want_tree else 0
        opts.rounds_per_sync = rounds_per_sync
        opts.chunk_arcs = chunk_arcs
        opts.pull_permille = pull_permille
        opts.grow_mode = grow_mode
        opts.push_lanes = push_lanes
        opts.half_bound = 0 if half_bound else -1
        opts.auto_bound = 1 if auto_bound else 0
        opts.tour_bound = 0 if tour_bound else -1
        opts.merge_tile_kb = merge_tile_kb
        res = _abi.gst_result()
        grp = groups.as_c()
        st = self.lib.gst_solve(self.ctx, graph.handle, C.byref(grp), C.byref(opts), C.byref(res))
        try:
            _check(self.lib, st, self.ctx)
            ne = res.n_edges
            if ne > 0:
                eu = np.ctypeslib.as_array(res.eu, shape=(ne,)).copy()
                ev = np.ctypeslib.as_array(res.ev, shape=(ne,)).copy()
                ew = np.ctypeslib.as_array(res.ew, shape=(ne,)).copy()
            else:
                eu = np.zeros(0, np.int32)
                ev = np.zeros(0, np.int32)
                ew = np.zeros(0, np.float64)
            stats = {name: getattr(res.stats, name) for name, _ in _abi.gst_stats._fields_}
            return Result(res.cost, res.tree_cost, res.root, eu, ev, ew, bool(res.used_ub), bool(res.fell_back),
                          stats)
        finally:
            self.lib.gst_result_free(C.byref(res))

    def mst_postprocess(self, graph: DeviceGraph, vertices) -> Tuple[np.ndarray, np.ndarray, np.ndarray, float]:
        """The minimum spanning forest of the subgraph induced by `vertices` (gst_mst_postprocess; the
        reference's graph_hash_of_mixed_weighted_MST_postprocessing_hash_of_vertices): (eu, ev, ew, cost)."""
        vs = np.ascontiguousarray(np.asarray(vertices, dtype=np.int32))
        n = max(int(vs.size), 1)
        eu = np.zeros(n, np.int32)
        ev = np.zeros(n, np.int32)
        ew = np.zeros(n, np.float64)
        ne = C.c_int32(0)
        cost = C.c_double(0.0)
        st = self.lib.gst_mst_postprocess(self.ctx, graph.handle, _ptr(vs), int(vs.size), _ptr(eu), _ptr(ev), _ptr(ew),
                                          C.byref(ne), C.byref(cost))
        _check(self.lib, st, self.ctx)
        k = ne.value
        return eu[:k].copy(), ev[:k].copy(), ew[:k].copy(), cost.value

    def solve_noalloc(self, graph: DeviceGraph, groups: Groups, *, ub: Optional[float] = None, prune_lb: bool = True,
                      want_tree: bool = True, edge_cap: int = 16384, auto_bound: bool = False, half_bound: bool = True,
                      tour_bound: bool = True, push_lanes: int = 0, pull_permille: int = 0) -> Result:
        """One exact query through gst_solve_noalloc: no allocation inside the call (the device
        state lives in the library's fixed scratch array); the tree goes into arrays made here."""
        opts = _abi.gst_opts()
        opts.ub = float(ub) if ub is not None else 0.0
        opts.prune_lb = 1 if prune_lb else 0
        opts.want_tree = 1 if want_tree else 0
        opts.push_lanes = push_lanes
        opts.pull_permille = pull_permille
        opts.half_bound = 0 if half_bound else -1
        opts.auto_bound = 1 if auto_bound else 0
        opts.tour_bound = 0 if tour_bound else -1
        eu = np.zeros(max(edge_cap, 1), np.int32)
        ev = np.zeros(max(edge_cap, 1), np.int32)
        ew = np.zeros(max(edge_cap, 1), np.float64)
        res = _abi.gst_result()
        grp = groups.as_c()
        st = self.lib.gst_solve_noalloc(self.ctx, graph.handle, C.byref(grp), C.byref(opts), int(edge_cap), _ptr(eu),
                                        _ptr(ev), _ptr(ew), C.byref(res))
        _check(self.lib, st, self.ctx)
        ne = res.n_edges
        stats = {name: getattr(res.stats, name) for name, _ in _abi.gst_stats._fields_}
        return Result(res.cost, res.tree_cost, res.root, eu[:ne].copy(), ev[:ne].copy(), ew[:ne].copy(),
                      bool(res.used_ub), bool(res.fell_back), stats)

    def solve_batch(self, graph: DeviceGraph, queries: List[Groups], *, ubs: Optional[Sequence[float]] = None,
                    prune_lb: bool = True, want_tree: bool = True, slots: int = 0, rounds_per_step: int = 0,
                    use_graph: bool = True, push_lanes: int = 0, half_bound: bool = True,
                    auto_bound: bool = False, pull_permille: int = 0, time_kernels: bool = False, tour_bound: bool = True,
                    raise_on_error: bool = True, max_steps: int = 0):
        """Solve many queries on one graph with the batched engine (gst_solve_batch).

        Returns (results, stats): results[i] is a Result, or a GstError for a query that failed
        when raise_on_error is False."""
        nq = len(queries)
        grps = (_abi.gst_groups * max(nq, 1))()
        keep = [q.as_c() for q in queries]
        for i, c in enumerate(keep):
            grps[i] = c
        opts = (_abi.gst_opts * max(nq, 1))()
        for i in range(nq):
            ub = ubs[i] if ubs is not None else None
            opts[i].ub = float(ub) if ub is not None else 0.0
            opts[i].prune_lb = 1 if prune_lb else 0
            opts[i].want_tree = 1 if want_tree else 0
            opts[i].push_lanes = push_lanes
            opts[i].half_bound = 0 if half_bound else -1
            opts[i].auto_bound = 1 if auto_bound else 0
            opts[i].tour_bound = 0 if tour_bound else -1
            opts[i].pull_permille = pull_permille
        bo = _abi.gst_batch_opts()
        bo.slots = slots
        bo.rounds_per_step = rounds_per_step
        bo.use_graph = 1 if use_graph else 0
        bo.time_kernels = 1 if time_kernels else 0
        bo.max_steps = max_steps
        res = (_abi.gst_result * max(nq, 1))()
        status = (C.c_int * max(nq, 1))()
        bs = _abi.gst_batch_stats()
        st = self.lib.gst_solve_batch(self.ctx, graph.handle, nq, grps, opts, C.byref(bo), res, status, C.byref(bs))
        out = []
        try:
            for i in range(nq):
                if status[i] != _abi.GST_OK:
                    out.append(GstError(status[i], "query %d failed" % i))
                    continue
                r = res[i]
                ne = r.n_edges
                if ne > 0:
                    eu = np.ctypeslib.as_array(r.eu, shape=(ne,)).copy()
                    ev = np.ctypeslib.as_array(r.ev, shape=(ne,)).copy()
                    ew = np.ctypeslib.as_array(r.ew, shape=(ne,)).copy()
                else:
                    eu, ev, ew = np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float64)
                stats = {name: getattr(r.stats, name) for name, _ in _abi.gst_stats._fields_}
                out.append(Result(r.cost, r.tree_cost, r.root, eu, ev, ew, bool(r.used_ub), bool(r.fell_back), stats))
        finally:
            for i in range(nq):
                self.lib.gst_result_free(C.byref(res[i]))
        stats = {name: getattr(bs, name) for name, _ in _abi.gst_batch_stats._fields_ if name != "reserved"}
        if st not in (_abi.GST_OK,) and raise_on_error:
            bad = next((e for e in out if isinstance(e, GstError)), None)
            raise GstError(st, self.lib.gst_last_error(self.ctx).decode()) if bad is None else bad
        return out, stats


class Model:
    """Learned-bound network on one Solver's device (see include/gst.h, gst_model_*)."""

    def __init__(self, solver: "Solver", seed: int = 0, params: Optional[np.ndarray] = None):
        self.solver = solver
        self.lib = solver.lib
        h = C.c_void_p()
        _check(self.lib, self.lib.gst_model_create(solver.ctx, seed & (2**64 - 1), C.byref(h)), solver.ctx)
        self.handle = h.value
        self.num_params = int(self.lib.gst_model_num_params())
        if params is not None:
            self.set_params(params)

    def free(self):
        if self.handle:
            self.lib.gst_model_free(self.handle)
            self.handle = None

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        _check(self.lib, self.lib.gst_model_get_params(self.handle, _ptr(out)), self.solver.ctx)
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.num_params,)
        _check(self.lib, self.lib.gst_model_set_params(self.handle, _ptr(p)), self.solver.ctx)

    def scores(self, graph: "DeviceGraph", groups: Groups) -> np.ndarray:
        out = np.zeros(graph.n, np.float32)
        grp = groups.as_c()
        _check(self.lib, self.lib.gst_model_scores(self.handle, graph.handle, C.byref(grp), _ptr(out)), self.solver.ctx)
        return out

    def loss_grad(self, graph: "DeviceGraph", groups: Groups, labels, pos_weight: float = 1.0, want_grad=False):
        labels = np.ascontiguousarray(labels, dtype=np.uint8)
        grad = np.zeros(self.num_params, np.float32) if want_grad else None
        loss = C.c_float(0)
        grp = groups.as_c()
        _check(self.lib, self.lib.gst_model_loss_grad(self.handle, graph.handle, C.byref(grp), _ptr(labels),
                                                      float(pos_weight), C.byref(loss),
                                                      _ptr(grad) if grad is not None else None), self.solver.ctx)
        return loss.value, grad

    def adam_step(self, lr: float) -> None:
        _check(self.lib, self.lib.gst_model_adam_step(self.handle, float(lr)), self.solver.ctx)

    def save(self, path: str) -> None:
        np.save(path, self.get_params())

    def load(self, path: str) -> None:
        self.set_params(np.load(path, allow_pickle=False))


def _solve_learned(self, graph: DeviceGraph, groups: Groups, model: Model, topk_frac: float = 0.05, *,
                   prune_lb: bool = True, want_tree: bool = True) -> Result:
    opts = _abi.gst_opts()
    opts.prune_lb = 1 if prune_lb else 0
    opts.want_tree = 1 if want_tree else 0
    res = _abi.gst_result()
    ls = _abi.gst_learned_stats()
    grp = groups.as_c()
    st = self.lib.gst_solve_learned(self.ctx, graph.handle, C.byref(grp), model.handle, float(topk_frac),
                                    C.byref(opts), C.byref(res), C.byref(ls))
    try:
        _check(self.lib, st, self.ctx)
        ne = res.n_edges
        if ne > 0:
            eu = np.ctypeslib.as_array(res.eu, shape=(ne,)).copy()
            ev = np.ctypeslib.as_array(res.ev, shape=(ne,)).copy()
            ew = np.ctypeslib.as_array(res.ew, shape=(ne,)).copy()
        else:
            eu, ev, ew = np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float64)
        stats = {name: getattr(res.stats, name) for name, _ in _abi.gst_stats._fields_}
        stats.update({name: getattr(ls, name) for name, _ in _abi.gst_learned_stats._fields_})
        return Result(res.cost, res.tree_cost, res.root, eu, ev, ew, bool(res.used_ub), bool(res.fell_back), stats)
    finally:
        self.lib.gst_result_free(C.byref(res))


def _solve_batch_learned(self, graph: DeviceGraph, queries: List[Groups], model: Model, topk_frac: float = 0.05, *,
                         prune_lb: bool = True, want_tree: bool = True, slots: int = 0, rounds_per_step: int = 0,
                         use_graph: bool = True, push_lanes: int = 0, half_bound: bool = True,
                         pull_permille: int = 0, time_kernels: bool = False, raise_on_error: bool = True,
                         tour_bound: bool = True, max_steps: int = 0):
    """Exact solves of many queries with the learned bound, pipelined through the batched engine
    (gst_solve_batch_learned).  Returns (results, batch stats); each Result's stats also carry the
    query's learned-bound fields (sub_vertices, ub, t_sub_solve_ms, ...)."""
    nq = len(queries)
    grps = (_abi.gst_groups * max(nq, 1))()
    keep = [q.as_c() for q in queries]
    for i, c in enumerate(keep):
        grps[i] = c
    opts = (_abi.gst_opts * max(nq, 1))()
    for i in range(nq):
        opts[i].prune_lb = 1 if prune_lb else 0
        opts[i].want_tree = 1 if want_tree else 0
        opts[i].push_lanes = push_lanes
        opts[i].half_bound = 0 if half_bound else -1
        opts[i].tour_bound = 0 if tour_bound else -1
        opts[i].pull_permille = pull_permille
    bo = _abi.gst_batch_opts()
    bo.slots = slots
    bo.rounds_per_step = rounds_per_step
    bo.use_graph = 1 if use_graph else 0
    bo.time_kernels = 1 if time_kernels else 0
    bo.max_steps = max_steps
    res = (_abi.gst_result * max(nq, 1))()
    status = (C.c_int * max(nq, 1))()
    ls = (_abi.gst_learned_stats * max(nq, 1))()
    bs = _abi.gst_batch_stats()
    st = self.lib.gst_solve_batch_learned(self.ctx, graph.handle, nq, grps, model.handle, float(topk_frac), opts,
                                          C.byref(bo), res, status, ls, C.byref(bs))
    out = []
    try:
        for i in range(nq):
            if status[i] != _abi.GST_OK:
                out.append(GstError(status[i], "query %d failed" % i))
                continue
            r = res[i]
            ne = r.n_edges
            if ne > 0:
                eu = np.ctypeslib.as_array(r.eu, shape=(ne,)).copy()
                ev = np.ctypeslib.as_array(r.ev, shape=(ne,)).copy()
                ew = np.ctypeslib.as_array(r.ew, shape=(ne,)).copy()
            else:
                eu, ev, ew = np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float64)
            stats = {name: getattr(r.stats, name) for name, _ in _abi.gst_stats._fields_}
            stats.update({name: getattr(ls[i], name) for name, _ in _abi.gst_learned_stats._fields_})
            out.append(Result(r.cost, r.tree_cost, r.root, eu, ev, ew, bool(r.used_ub), bool(r.fell_back), stats))
    finally:
        for i in range(nq):
            self.lib.gst_result_free(C.byref(res[i]))
    stats = {name: getattr(bs, name) for name, _ in _abi.gst_batch_stats._fields_ if name != "reserved"}
    if st != _abi.GST_OK and raise_on_error:
        bad = next((e for e in out if isinstance(e, GstError)), None)
        raise GstError(st, self.lib.gst_last_error(self.ctx).decode()) if bad is None else bad
    return out, stats


def _autotune(self, graph: DeviceGraph, sample: List[Groups], *, bounded: bool = False, batch: bool = False,
              eager: bool = False) -> dict:
    """Pick the launch knobs for this graph from a sample of queries (gst_autotune); later solves
    that leave a knob at 0 use the choice.  Returns the tuning report as a dict."""
    nq = len(sample)
    grps = (_abi.gst_groups * max(nq, 1))()
    keep = [q.as_c() for q in sample]
    for i, c in enumerate(keep):
        grps[i] = c
    flags = (_abi.GST_TUNE_BOUNDED if bounded else 0) | (_abi.GST_TUNE_BATCH if batch else 0) | \
        (_abi.GST_TUNE_EAGER if eager else 0)
    rep = _abi.gst_tune_report()
    _check(self.lib, self.lib.gst_autotune(self.ctx, graph.handle, nq, grps, flags, C.byref(rep)), self.ctx)
    return {name: getattr(rep, name) for name, _ in _abi.gst_tune_report._fields_ if name != "pad"}


Solver.solve_learned = _solve_learned
Solver.autotune = _autotune
Solver.solve_batch_learned = _solve_batch_learned
__all__ += ["Model"]
