"""MST post-processing on the device (gst_mst_postprocess and the spider tree of the automatic bound)
against the oracle's restatement of the reference's post-processing."""
import numpy as np
import pytest

import gpu4gst_code_amd as gst
import oracle

pytestmark = pytest.mark.gpu


def _edge_weights(G):
    d = {}
    for v in range(G.n):
        for e in range(int(G.row_ptr[v]), int(G.row_ptr[v + 1])):
            x = int(G.col[e])
            key = (min(v, x), max(v, x))
            d[key] = min(d.get(key, np.inf), float(G.w[e]))
    return d


def _check_forest(G, vs, eu, ev, ew, wmap):
    inset = set(int(v) for v in vs)
    par = {v: v for v in inset}

    def root(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x
    for u, v, w in zip(eu, ev, ew):
        u, v = int(u), int(v)
        assert u in inset and v in inset and u != v
        assert wmap[(min(u, v), max(u, v))] == w
        ru, rv = root(u), root(v)
        assert ru != rv, "cycle"
        par[ru] = rv


def _ball(G, src, size):
    seen, order = {src}, [src]
    for v in order:
        for e in range(int(G.row_ptr[v]), int(G.row_ptr[v + 1])):
            x = int(G.col[e])
            if x not in seen and len(order) < size:
                seen.add(x)
                order.append(x)
    return order


def test_mst_postprocess_int_vs_oracle(solver):
    """Connected sets (BFS balls) and random, mostly disconnected, sets on a V=2000 graph: the forest
    has the oracle's cost exactly, its edges are graph edges with their weights, it has no cycle and
    the oracle's edge count."""
    G = gst.gen_random_connected(2000, 10000, 1, 100, seed=91)
    dg = solver.upload(G)
    wmap = _edge_weights(G)
    rng = np.random.default_rng(3)
    sets = [_ball(G, int(rng.integers(0, G.n)), int(s)) for s in (2, 10, 60, 200, 500)]
    sets += [rng.integers(0, G.n, size=int(s)) for s in (1, 5, 100, 450)]
    sets.append(np.concatenate([sets[2], sets[2]]))   # repeats are ignored
    for vs in sets:
        eu, ev, ew, cost = solver.mst_postprocess(dg, vs)
        want_edges, want = oracle.mst_induced(G, vs)
        print(f"|set| {len(set(int(x) for x in vs))}: gpu {cost} ({len(eu)} edges), oracle {want} ({len(want_edges)})",
              flush=True)
        assert cost == want
        assert len(eu) == len(want_edges)
        _check_forest(G, vs, eu, ev, ew, wmap)
    dg.free()


def test_mst_postprocess_float_and_limits(solver):
    G = gst.gen_rmat(12, 8, weights="lognormal", seed=21)
    dg = solver.upload(G)
    wmap = _edge_weights(G)
    rng = np.random.default_rng(4)
    for s in (3, 40, 300):
        vs = _ball(G, int(np.argmax(np.diff(G.row_ptr))), s)
        eu, ev, ew, cost = solver.mst_postprocess(dg, vs)
        _, want = oracle.mst_induced(G, vs)
        assert abs(cost - want) <= 1e-6 * max(1.0, want)
        _check_forest(G, vs, eu, ev, ew, wmap)
    with pytest.raises(gst.GstError):
        solver.mst_postprocess(dg, np.arange(600))          # more than 512 distinct vertices
    with pytest.raises(gst.GstError):
        solver.mst_postprocess(dg, [0, G.n])                 # out of range
    _ = rng
    dg.free()


@pytest.mark.parametrize("noalloc", [False, True])
def test_spider_tree_tightens_the_automatic_bound(solver, noalloc):
    """The automatic bound's spider (min_c sum_i d(c, G_i), equal to the oracle's) is post-processed
    into a tree whose cost lies between the optimum and the spider, strictly below the spider on
    some queries, and every solve keeps the exact optimum."""
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    dg = solver.upload(G)
    tighter = 0
    for q in range(10):
        groups = gst.gen_groups(G.n, 6, 10, 200, seed=2000 + q)
        r = solver.solve_noalloc(dg, groups, auto_bound=True) if noalloc else solver.solve(dg, groups, auto_bound=True)
        opt = oracle.closed_form(G, groups) if groups.g <= 4 else solver.solve(dg, groups).cost
        sc, tc = r.stats["spider_cost"], r.stats["spider_tree_cost"]
        print(f"q{q}: opt {opt} spider {sc} tree {tc} cost {r.cost}", flush=True)
        assert sc == oracle.spider_cost(G, groups)
        assert opt <= tc <= sc
        assert r.cost == opt
        tighter += tc < sc
    assert tighter > 0
    dg.free()


def test_spider_tree_in_batches_and_floats(solver):
    """Batched spider-bounded solves (their level-2 merges find the spider, a workgroup per slot
    post-processes it) equal the unbounded costs; float weights stay within 1e-6 of the oracle."""
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    dg = solver.upload(G)
    qs = [gst.gen_groups(G.n, int(g), 10, 200, seed=2100 + i) for i, g in enumerate([3, 4, 5, 6, 7, 6, 5, 8])]
    res, _ = solver.solve_batch(dg, qs, auto_bound=True, slots=4)
    for q, r in zip(qs, res):
        assert r.cost == solver.solve(dg, q).cost
    dg.free()
    Gf = gst.gen_rmat(12, 8, weights="lognormal", seed=21)
    dgf = solver.upload(Gf)
    for i in range(4):
        q = gst.gen_groups(Gf.n, 5, 2, 30, seed=60 + i)
        want = oracle.dpbf_cost(Gf, q)
        r = solver.solve(dgf, q, auto_bound=True)
        assert abs(r.cost - want) <= 1e-6 * max(1.0, want)
        assert r.stats["spider_tree_cost"] <= r.stats["spider_cost"] * (1 + 1e-6)
        assert r.stats["spider_tree_cost"] >= want * (1 - 1e-6)
    dgf.free()
