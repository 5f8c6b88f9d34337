"""CPU: pin the oracle (label-setting DPBF restatement) against brute force, closed forms and goldens."""
import json
import os

import numpy as np
import pytest

import gpu4gst_code_amd as gst
import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


@pytest.mark.parametrize("s", range(0, 120, 7))
def test_tiny_vs_bruteforce(s):
    n = 10 + s % 5
    g = 1 + s % 6
    G = gst.gen_random_connected(n, n + 6, 1, 20, seed=5000 + s)
    groups = gst.gen_groups(n, g, 1, 3, seed=6000 + s)
    c = oracle.dpbf_cost(G, groups)
    assert c == oracle.bruteforce(G, groups)
    assert c == GOLD["tiny"]["costs"][str(s)]


@pytest.mark.parametrize("s", range(40))
def test_float_weights_vs_bruteforce(s):
    """Float weights: oracle (double) equals brute force up to summation order."""
    rng = np.random.default_rng(s)
    n = 12
    G0 = gst.gen_random_connected(n, 20, 1, 1, seed=700 + s)
    w = rng.lognormal(0.0, 1.0, size=G0.m // 2).astype(np.float32)
    src = np.repeat(np.arange(n), np.diff(G0.row_ptr))
    keep = src < G0.col
    G = gst.HostGraph.from_edges(n, src[keep], G0.col[keep], w)
    groups = gst.gen_groups(n, 1 + s % 5, 1, 3, seed=800 + s)
    c = oracle.dpbf_cost(G, groups)
    b = oracle.bruteforce(G, groups)
    assert abs(c - b) <= 1e-12 * max(1.0, b)


@pytest.mark.parametrize("g", [1, 2, 3, 4])
@pytest.mark.parametrize("s", range(4))
def test_closed_forms(g, s):
    G = gst.gen_random_connected(1000, 5000, 1, 100, seed=40 + s)
    groups = gst.gen_groups(G.n, g, 5, 20, seed=50 + s + 10 * g)
    assert oracle.dpbf_cost(G, groups) == oracle.closed_form(G, groups)


def test_closed_form_g4_ba():
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=3)
    for q in range(2):
        groups = gst.gen_groups(G.n, 4, 10, 200, seed=90 + q)
        assert oracle.dpbf_cost(G, groups) == oracle.closed_form(G, groups)


def test_c1_goldens():
    for s in ("0", "1", "2", "17", "199"):
        G = gst.gen_random_connected(1000, 5000, 1, 100, seed=int(s))
        groups = gst.gen_groups(G.n, 3, 5, 20, seed=int(s) + 1)
        assert oracle.dpbf_cost(G, groups) == GOLD["c1"]["costs"][s]


def test_reference_fixture():
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "gstp_fixture.npz"))
    G = gst.HostGraph(z["row_ptr"], z["col"], z["w"])
    groups = gst.Groups(z["grp_ptr"], z["grp_vtx"])
    c, root, eu, ev, ew, _ = oracle.dpbf(G, groups)
    assert c == GOLD["gstp_fixture"]["cost"] == 16.0
    assert oracle.closed_form(G, groups) == c
    assert gst.validate_tree(G, groups, eu, ev, root) == c


def test_oracle_tree_is_valid_and_costs_match():
    for s in range(5):
        G = gst.gen_random_connected(1000, 5000, 1, 100, seed=300 + s)
        groups = gst.gen_groups(G.n, 5, 5, 20, seed=400 + s)
        c, root, eu, ev, ew, _ = oracle.dpbf(G, groups)
        assert gst.validate_tree(G, groups, eu, ev, root) == c
        assert float(np.sum(ew)) == c


def test_infeasible():
    # two components: groups on different sides
    G = gst.HostGraph.from_edges(6, [0, 1, 3, 4], [1, 2, 4, 5], np.array([1, 1, 1, 1], np.int32))
    groups = gst.groups_from_lists([[0], [5]])
    assert oracle.dpbf_cost(G, groups) == np.inf
    assert oracle.bruteforce(G, groups) == np.inf


def test_planted_mistakes_are_caught():
    """A plausible oracle mistake must fail a pin: dropping a group bit (wrong member mapping)
    or counting an edge twice changes the cost seen by the closed form / brute force."""
    G = gst.gen_random_connected(1000, 5000, 1, 100, seed=1)
    groups = gst.gen_groups(G.n, 3, 5, 20, seed=2)
    good = oracle.dpbf_cost(G, groups)
    # mistake 1: group 2 loses its first member
    lists = [list(groups.members(i)) for i in range(3)]
    bad_groups = gst.groups_from_lists([lists[0], lists[1], lists[2][1:]])
    # the pin recomputed on the true groups disagrees unless the dropped member was irrelevant
    diffs = 0
    for drop in range(len(lists[2])):
        bg = gst.groups_from_lists([lists[0], lists[1], lists[2][:drop] + lists[2][drop + 1:]])
        diffs += oracle.dpbf_cost(G, bg) != oracle.closed_form(G, groups)
    assert diffs >= 1
    # mistake 2: doubled weights on the traced tree break the cost identity
    c, root, eu, ev, ew, _ = oracle.dpbf(G, groups)
    assert 2 * float(np.sum(ew)) != good
    _ = bad_groups


def _zero_weight_graph(n, extra, whi, seed):
    return gst.gen_random_connected(n, n + extra, 0, whi, seed=seed)


@pytest.mark.parametrize("s", range(0, 60, 3))
def test_pruned_half_bound_vs_bruteforce(s):
    """Bounds of PrunedDP++ (keep cost + LB <= U, grow only cost <= U/2) keep the optimum for any
    U >= OPT: tiny graphs against brute force, with zero-weight edges on every third seed."""
    n = 10 + s % 5
    g = 1 + s % 6
    G = _zero_weight_graph(n, 8, 4, 9000 + s) if s % 3 == 0 else gst.gen_random_connected(n, n + 6, 1, 20, seed=5000 + s)
    groups = gst.gen_groups(n, g, 1, 3, seed=6000 + s)
    b = oracle.bruteforce(G, groups)
    for U in (b, b + 1, 2 * b + 3):
        for lb in (True, False):
            c, _ = oracle.dpbf_pruned(G, groups, U, 0.5, lb)
            assert c == b, (U, lb, c, b)


@pytest.mark.parametrize("s", range(6))
def test_pruned_half_bound_c1_and_float(s):
    G = gst.gen_random_connected(1000, 5000, 1, 100, seed=s)
    groups = gst.gen_groups(G.n, 3 + s % 4, 5, 20, seed=s + 1)
    c0, st0 = oracle.dpbf(G, groups)[0], oracle.dpbf(G, groups)[5]
    for U in (c0, c0 * 1.1, c0 * 3):
        c, st = oracle.dpbf_pruned(G, groups, U)
        assert c == c0
        # label-setting stops at OPT, so a loose bound (U >= 2 OPT) may prune nothing
        assert st["relax"] <= st0["relax"]
        if U == c0:
            assert st["relax"] < oracle.dpbf_pruned(G, groups, U, 0.0)[1]["relax"]
    # a bound below the optimum leaves no tree
    assert oracle.dpbf_pruned(G, groups, c0 - 1)[0] == np.inf
    Gf = gst.gen_rmat(10, 8, weights="lognormal", seed=40 + s)
    gf = gst.gen_groups(Gf.n, 4, 3, 20, seed=50 + s)
    cf = oracle.dpbf_cost(Gf, gf)
    assert abs(oracle.dpbf_pruned(Gf, gf, cf * (1 + 1e-9))[0] - cf) <= 1e-12 * cf


def test_pruned_grow_limit_below_half_is_caught():
    """Planted mistake: growing only states of cost <= U/3 loses the optimum on some instance,
    so the tests above would catch a wrong grow limit."""
    lost = 0
    for s in range(40):
        G = gst.gen_random_connected(12, 18, 1, 20, seed=7000 + s)
        groups = gst.gen_groups(12, 2 + s % 4, 1, 2, seed=7100 + s)
        b = oracle.bruteforce(G, groups)
        lost += oracle.dpbf_pruned(G, groups, b, 1.0 / 3.0)[0] != b
    assert lost >= 1


@pytest.mark.parametrize("s", range(8))
def test_spider_bound_is_feasible_and_keeps_optimum(s):
    """The automatic bound min_c sum_i dist(c, G_i) is the cost of a connected subgraph touching
    every group, so it is >= OPT (= OPT for g <= 3), and the bounded search under it keeps OPT."""
    G = gst.gen_random_connected(1000, 5000, 1, 100, seed=60 + s) if s % 2 == 0 else \
        gst.gen_random_connected(600, 2400, 0, 5, seed=60 + s)
    g = 3 + s % 5
    groups = gst.gen_groups(G.n, g, 3, 20, seed=70 + s)
    opt = oracle.dpbf_cost(G, groups)
    spider = float(np.min(oracle.group_distances(G, groups).sum(axis=0)))
    assert spider >= opt
    if g == 3:
        assert spider == opt
    assert oracle.dpbf_pruned(G, groups, spider)[0] == opt


@pytest.mark.parametrize("s", range(0, 60, 3))
def test_tour_bound_keeps_optimum_vs_bruteforce(s):
    """One-label plus the tour bounds of PrunedDP++ keep the brute-force optimum for U >= OPT."""
    n = 10 + s % 5
    g = 2 + s % 5
    G = _zero_weight_graph(n, 8, 4, 9100 + s) if s % 3 == 0 else gst.gen_random_connected(n, n + 6, 1, 20, seed=5100 + s)
    groups = gst.gen_groups(n, g, 1, 3, seed=6100 + s)
    b = oracle.bruteforce(G, groups)
    for U in (b, b + 1, 2 * b + 3):
        assert oracle.dpbf_pruned(G, groups, U, 0.5, 2)[0] == b
        assert oracle.dpbf_pruned(G, groups, U, 0.0, 2)[0] == b


@pytest.mark.parametrize("s", range(0, 60, 4))
def test_pop_upper_bound_keeps_optimum_vs_bruteforce(s):
    """PrunedDP++'s feasible tree at every pop (popped tree + shortest paths to the missing
    groups) and its exact early exit keep the brute-force optimum for every U >= OPT, from an
    unbounded start too (U = inf: the pops alone supply the bound), zero weights included."""
    n = 10 + s % 5
    g = 2 + s % 5
    G = _zero_weight_graph(n, 8, 4, 9300 + s) if s % 3 == 0 else gst.gen_random_connected(n, n + 6, 1, 20, seed=5300 + s)
    groups = gst.gen_groups(n, g, 1, 3, seed=6300 + s)
    b = oracle.bruteforce(G, groups)
    for U in (b, b + 1, 2 * b + 3, np.inf):
        for lb in (1, 2):
            for frac in (0.5, 0.0):
                c, _ = oracle.dpbf_pruned(G, groups, U, frac if np.isfinite(U) else 0.0, lb, pop_ub=True)
                assert c == b, (U, lb, frac, c, b)


def test_pop_upper_bound_prunes_on_c1():
    """On C1-sized instances the per-pop bound and early exit pop fewer states than the bounded
    search without them, at the same optimum."""
    fewer = 0
    for s in range(4):
        G = gst.gen_random_connected(1000, 5000, 1, 100, seed=s)
        groups = gst.gen_groups(G.n, 5, 5, 20, seed=s + 1)
        c0 = oracle.dpbf_cost(G, groups)
        d = oracle.group_distances(G, groups)
        spider = float(np.min(d.sum(axis=0)))
        c1, st1 = oracle.dpbf_pruned(G, groups, spider, 0.5, 2, dist=d)
        c2, st2 = oracle.dpbf_pruned(G, groups, spider, 0.5, 2, pop_ub=True, dist=d)
        assert c1 == c0 and c2 == c0
        fewer += st2["pops"] < st1["pops"]
    assert fewer >= 1


@pytest.mark.parametrize("s", range(6))
def test_tour_bound_per_state_vs_bruteforce(s):
    """Each state's bound is at most the cost of the cheapest tree that holds v and touches every
    group missing from the state (brute force with {v} as an extra group), and is never below the
    one-label bound; with one group missing the two bounds agree."""
    n = 9 + s % 3
    g = 3 + s % 2
    G = gst.gen_random_connected(n, n + 5, 1, 12, seed=8100 + s) if s % 2 else _zero_weight_graph(n, 6, 3, 8200 + s)
    groups = gst.gen_groups(n, g, 1, 3, seed=8300 + s)
    lists = [list(groups.members(i)) for i in range(g)]
    tour = oracle.state_bounds(G, groups, 2)
    one = oracle.state_bounds(G, groups, 1)
    full = (1 << g) - 1
    for p in range(1, full):
        miss = [i for i in range(g) if not (p >> i) & 1]
        for v in range(n):
            need = oracle.bruteforce(G, gst.groups_from_lists([lists[i] for i in miss] + [[v]]))
            assert one[v, p] <= tour[v, p] <= need + 1e-12, (v, p, tour[v, p], need)
            if len(miss) == 1:
                assert tour[v, p] == one[v, p]


def test_tour_bound_is_stronger_somewhere():
    strict = 0
    for s in range(6):
        G = gst.gen_random_connected(200, 600, 1, 50, seed=8400 + s)
        groups = gst.gen_groups(G.n, 5, 2, 6, seed=8500 + s)
        strict += int(np.sum(oracle.state_bounds(G, groups, 2) > oracle.state_bounds(G, groups, 1)))
    assert strict > 0


def test_tour_bound_planted_mistake_is_caught():
    """Planted mistake: taking lb1 as the max (not the min) over the tour's end groups overstates
    the bound and loses the optimum on some instance, so the tests above would catch it."""
    lost = 0
    for s in range(30):
        G = gst.gen_random_connected(12, 18, 1, 20, seed=7200 + s)
        groups = gst.gen_groups(12, 3 + s % 3, 1, 2, seed=7300 + s)
        b = oracle.bruteforce(G, groups)
        lost += oracle.dpbf_pruned(G, groups, b, 0.5, 3)[0] != b
    assert lost >= 1


@pytest.mark.parametrize("s", range(4))
def test_tour_bound_prunes_more_c1(s):
    G = gst.gen_random_connected(1000, 5000, 1, 100, seed=3 + s)
    groups = gst.gen_groups(G.n, 6, 5, 20, seed=s + 1)
    c0 = oracle.dpbf_cost(G, groups)
    c1, st1 = oracle.dpbf_pruned(G, groups, c0, 0.5, 1)
    c2, st2 = oracle.dpbf_pruned(G, groups, c0, 0.5, 2)
    assert c1 == c2 == c0
    assert st2["pops"] <= st1["pops"]


def _brute_msf_cost(n, edges):
    """Minimum spanning forest cost by enumerating edge subsets (tiny graphs only)."""
    import itertools
    # components of the whole edge set fix the forest size
    par = list(range(n))

    def root(x):
        while par[x] != x:
            x = par[x]
        return x
    for a, b, _ in edges:
        ra, rb = root(a), root(b)
        if ra != rb:
            par[ra] = rb
    k = n - len({root(x) for x in range(n)})
    best = None
    for sub in itertools.combinations(edges, k):
        p2 = list(range(n))

        def r2(x):
            while p2[x] != x:
                x = p2[x]
            return x
        ok = True
        for a, b, _ in sub:
            ra, rb = r2(a), r2(b)
            if ra == rb:
                ok = False
                break
            p2[ra] = rb
        if ok:
            c = sum(wt for _, _, wt in sub)
            best = c if best is None else min(best, c)
    return best if best is not None else 0.0


def test_mst_induced_pinned():
    """oracle.mst_induced (the reference's MST post-processing) against brute force over every
    edge subset on tiny graphs and against scipy's minimum spanning tree on larger induced sets,
    disconnected sets included.  A planted mistake (the same Prim taking the heaviest edge: a maximum
    spanning forest) fails the brute-force check on some instance, so the check can see it."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import minimum_spanning_tree

    class Neg:   # the graph with negated weights: Prim on it is the planted maximum spanning forest
        def __init__(self, G):
            self.row_ptr, self.col, self.w = G.row_ptr, G.col, -np.asarray(G.w, dtype=np.float64)

    rng = np.random.default_rng(5)
    planted_caught = 0
    for t in range(40):
        n = int(rng.integers(3, 7))
        G = gst.gen_random_connected(n, min(n * (n - 1) // 2, n + int(rng.integers(1, 6))), 1, 9, seed=300 + t)
        vs = [v for v in range(n) if rng.random() < 0.85] or [0]
        edges, cost = oracle.mst_induced(G, vs)
        idx = {v: i for i, v in enumerate(vs)}
        ind = []
        for v in vs:
            for e in range(int(G.row_ptr[v]), int(G.row_ptr[v + 1])):
                x = int(G.col[e])
                if x in idx and v < x:
                    ind.append((idx[v], idx[x], float(G.w[e])))
        want = _brute_msf_cost(len(vs), ind)
        assert cost == want, (t, cost, want)
        assert all(u < v and u in idx and v in idx for u, v, _ in edges)
        _, neg = oracle.mst_induced(Neg(G), vs)
        planted_caught += (-neg != want)
    assert planted_caught > 0
    for t in range(10):
        G = gst.gen_random_connected(600, 3000, 1, 100, seed=400 + t)
        vs = np.unique(rng.integers(0, 600, size=int(rng.integers(20, 400))))
        edges, cost = oracle.mst_induced(G, vs)
        sub = {int(v): i for i, v in enumerate(vs)}
        rows, cols, ws = [], [], []
        for v in vs:
            for e in range(int(G.row_ptr[v]), int(G.row_ptr[v + 1])):
                x = int(G.col[e])
                if x in sub:
                    rows.append(sub[int(v)]); cols.append(sub[x]); ws.append(float(G.w[e]))
        M = sp.csr_matrix((ws, (rows, cols)), shape=(len(vs), len(vs)))
        assert abs(minimum_spanning_tree(M).sum() - cost) < 1e-9, t
