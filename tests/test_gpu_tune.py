"""GPU: the autotuner picks knobs from its candidate sets, never reports a choice slower than the
defaults it measured, and solves after tuning stay exact (one-query, batched and learned)."""
import json
import os

import pytest

import gpu4gst_code_amd as gst

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


@pytest.mark.parametrize("bounded", [False, True])
def test_autotune_then_exact(solver, bounded):
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    dg = solver.upload(G)
    keys = sorted(GOLD["ba20k"]["costs"], key=int)
    qs = [gst.gen_groups(G.n, 6, 10, 200, seed=1000 + int(k)) for k in keys]
    rep = solver.autotune(dg, qs[:4], bounded=bounded, batch=True)
    print(rep, flush=True)
    assert rep["pull_permille"] in (450, -1, 250, 650, 850)
    assert rep["chunk_arcs"] in (32, 16, 64)
    assert rep["push_lanes"] in (64, 16, 32)
    assert rep["merge_tile_kb"] in (16, 8, 32, 48)
    assert rep["batch_rounds_per_step"] in (4, 8, 16)
    assert rep["batch_push_lanes"] in (8, 16, 32)
    assert rep["tuned_ms"] <= rep["default_ms"]
    assert rep["batch_tuned_qps"] >= rep["batch_default_qps"]
    assert rep["configs_tried"] >= 10
    for k, q in zip(keys, qs):
        want = GOLD["ba20k"]["costs"][k]
        r = solver.solve(dg, q, ub=want * 1.05 if bounded else None)
        assert r.cost == want
        assert gst.validate_tree(G, q, r.eu, r.ev, r.root) == want
    res, _ = solver.solve_batch(dg, qs, slots=4)
    for k, r in zip(keys, res):
        assert r.cost == GOLD["ba20k"]["costs"][k]
    model = gst.Model(solver, seed=3)
    res, _ = solver.solve_batch_learned(dg, qs, model, 0.05, slots=4)
    for k, q, r in zip(keys, qs, res):
        assert r.cost == GOLD["ba20k"]["costs"][k]
        assert gst.validate_tree(G, q, r.eu, r.ev, r.root) == r.cost
    model.free()
    dg.free()


def test_autotune_rejects_mixed_group_counts(solver):
    G = gst.gen_random_connected(500, 2000, 1, 50, seed=5)
    dg = solver.upload(G)
    qs = [gst.gen_groups(G.n, 3, 2, 10, seed=1), gst.gen_groups(G.n, 4, 2, 10, seed=2)]
    with pytest.raises(gst.GstError) as e:
        solver.autotune(dg, qs)
    assert e.value.status == gst._abi.GST_INVALID_ARG
    dg.free()
