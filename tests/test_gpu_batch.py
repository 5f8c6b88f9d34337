"""Batched engine (gst_solve_batch): every query of a batch against the oracle, the one-query
solve and the committed golden costs, with slots refilled many times over."""
import json
import os

import numpy as np
import pytest

import gpu4gst_code_amd as gst
import oracle

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def _valid(G, groups, r, rel=0.0):
    tc = gst.validate_tree(G, groups, r.eu, r.ev, r.root)
    assert abs(tc - r.cost) <= rel * max(1.0, abs(r.cost))


@pytest.mark.parametrize("slots,rounds,graph", [(4, 4, True), (16, 2, True), (64, 8, True), (8, 4, False)])
def test_mixed_group_counts_vs_oracle(solver, slots, rounds, graph):
    """40 queries, g = 1..8 and uneven group sizes, on a V=1000 graph; far more queries than slots."""
    G = gst.gen_random_connected(1000, 5000, 1, 100, seed=3)
    dg = solver.upload(G)
    rng = np.random.default_rng(11)
    qs = []
    for i in range(40):
        g = int(rng.integers(1, 9))
        qs.append(gst.gen_groups(G.n, g, 1, int(rng.integers(2, 40)), seed=1000 + i))
    print(f"slots {slots} rounds {rounds} graph {graph}", flush=True)
    res, st = solver.solve_batch(dg, qs, slots=slots, rounds_per_step=rounds, use_graph=graph)
    assert st["queries"] == 40 and st["slots"] >= 1
    for q, r in zip(qs, res):
        want = oracle.dpbf_cost(G, q)
        assert r.cost == want, (q.g, r.cost, want)
        _valid(G, q, r)
    dg.free()


def test_batch_matches_single_and_goldens_ba20k(solver):
    """The BA V=20k golden queries (g=6) through the engine, mixed with bounded copies of them."""
    gd = GOLDEN["ba20k"]
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    dg = solver.upload(G)
    keys = sorted(gd["costs"], key=int)
    qs = [gst.gen_groups(G.n, 6, 10, 200, seed=1000 + int(k)) for k in keys]
    want = [gd["costs"][k] for k in keys]
    ubs = [None] * len(qs) + want
    res, st = solver.solve_batch(dg, qs + qs, ubs=ubs, slots=6)
    for i, r in enumerate(res):
        q = qs[i % len(qs)]
        assert r.cost == want[i % len(qs)]
        _valid(G, q, r)
        assert r.used_ub == (i >= len(qs))
    single = solver.solve(dg, qs[0])
    assert single.cost == res[0].cost
    dg.free()


def test_float_weights_batch(solver):
    """Float32 weights: the engine runs the one-query kernel bodies, so costs agree with gst_solve
    to within 1e-6 relative (and with the double-precision oracle)."""
    G = gst.gen_rmat(12, 8, weights="lognormal", seed=21)
    dg = solver.upload(G)
    qs = [gst.gen_groups(G.n, int(g), 2, 30, seed=50 + i) for i, g in enumerate([2, 3, 5, 7, 4, 6])]
    res, _ = solver.solve_batch(dg, qs, slots=3)
    for q, r in zip(qs, res):
        one = solver.solve(dg, q)
        assert abs(r.cost - one.cost) <= 1e-6 * max(1.0, one.cost)
        want = oracle.dpbf_cost(G, q)
        assert abs(r.cost - want) <= 1e-6 * max(1.0, want)
        _valid(G, q, r, rel=1e-6)
    dg.free()


def test_failures_are_per_query(solver):
    """An infeasible query (a group in another component), an empty group and a bound below the
    optimum fail on their own; the other queries of the batch are still solved."""
    G = gst.HostGraph.from_edges(6, [0, 1, 3, 4], [1, 2, 4, 5], np.array([1, 1, 1, 1], np.int32))
    dg = solver.upload(G)
    ok = gst.groups_from_lists([[0], [2]])
    far = gst.groups_from_lists([[0], [5]])
    empty = gst.groups_from_lists([[0], []])
    res, st = solver.solve_batch(dg, [ok, far, empty, ok, ok], ubs=[None, None, None, 1.0, 2.0],
                                 raise_on_error=False, slots=2)
    assert res[0].cost == 2.0
    assert isinstance(res[1], gst.GstError) and res[1].status == gst._abi.GST_INFEASIBLE
    assert isinstance(res[2], gst.GstError) and res[2].status == gst._abi.GST_INFEASIBLE
    assert isinstance(res[3], gst.GstError) and res[3].status == gst._abi.GST_INFEASIBLE
    assert res[4].cost == 2.0
    with pytest.raises(gst.GstError):
        solver.solve_batch(dg, [ok, far])
    dg.free()


def test_zero_weights_batch(solver):
    """Zero-weight edges (equal-cost predecessors) still trace to a tree in the engine."""
    G = gst.gen_random_connected(400, 1600, 0, 3, seed=8)
    dg = solver.upload(G)
    qs = [gst.gen_groups(G.n, 5, 2, 10, seed=70 + i) for i in range(10)]
    res, _ = solver.solve_batch(dg, qs, slots=5)
    for q, r in zip(qs, res):
        assert r.cost == oracle.dpbf_cost(G, q)
        _valid(G, q, r)
    dg.free()


def test_long_trees_batch(solver):
    """Trees longer than the edges read back with a finished query's header (1024) take a second
    copy: paths of 3,000 and 1,500 edges next to short trees in the same batch."""
    n = 3000
    src = np.arange(n - 1)
    G = gst.HostGraph.from_edges(n, src, src + 1, np.ones(n - 1, np.int32))
    dg = solver.upload(G)
    qs = [gst.groups_from_lists([[0], [n - 1]]), gst.groups_from_lists([[10], [1510], [700]]),
          gst.groups_from_lists([[5], [9]]), gst.groups_from_lists([[100], [101], [102]])]
    want = [n - 1, 1500, 4, 2]
    res, _ = solver.solve_batch(dg, qs, slots=2)
    for q, r, w in zip(qs, res, want):
        assert r.cost == w
        assert len(r.eu) == w
        assert gst.validate_tree(G, q, r.eu, r.ev, r.root) == w
    dg.free()


def test_default_slots_follow_query_count(solver):
    """Without a slot count the engine sizes no more slots than there are queries to hold
    (queries that fail validation hold none); an explicit count is kept.  The count reported is
    the number of slots the run used (an engine with more slots from an earlier run keeps them idle)."""
    G = gst.gen_random_connected(500, 2000, 1, 50, seed=4)
    dg = solver.upload(G)
    qs = [gst.gen_groups(G.n, 4, 2, 10, seed=90 + i) for i in range(3)]
    bad = gst.groups_from_lists([[0], []])
    res, st = solver.solve_batch(dg, qs + [bad], raise_on_error=False)
    assert st["slots"] == 3
    for q, r in zip(qs, res):
        assert r.cost == oracle.dpbf_cost(G, q)
    _, st = solver.solve_batch(dg, qs, slots=5)
    assert st["slots"] == 5
    dg.free()


@pytest.mark.parametrize("learned", [False, True])
def test_engine_stop_marks_every_unfinished_query(solver, learned):
    """A run stopped early by the engine (here its step limit, set from the host: every kernel still
    ends normally) reports an error for every query it did not finish, never a zeroed result marked
    OK; the next call on the same engine solves the same queries exactly."""
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    dg = solver.upload(G)
    qs = [gst.gen_groups(G.n, 6, 10, 200, seed=1000 + i) for i in range(12)]
    if learned:
        model = gst.Model(solver, seed=3)
        res, _ = solver.solve_batch_learned(dg, qs, model, 0.05, slots=4, max_steps=3, raise_on_error=False)
    else:
        res, _ = solver.solve_batch(dg, qs, slots=4, max_steps=3, raise_on_error=False)
    bad = [r for r in res if isinstance(r, gst.GstError)]
    print(f"learned={learned}: {len(bad)} of {len(qs)} queries stopped", flush=True)
    assert bad, "three steps cannot finish twelve g=6 queries on four slots"
    for q, r in zip(qs, res):
        if isinstance(r, gst.GstError):
            assert r.status == 8   # GST_INTERNAL
        else:
            assert r.root >= 0 and gst.validate_tree(G, q, r.eu, r.ev, r.root) == r.cost
    again, _ = (solver.solve_batch_learned(dg, qs, model, 0.05, slots=4) if learned
                else solver.solve_batch(dg, qs, slots=4))
    for q, r in zip(qs, again):
        assert r.cost == solver.solve(dg, q, want_tree=False).cost
        _valid(G, q, r)
    if learned:
        model.free()
    dg.free()
