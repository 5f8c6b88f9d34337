"""GPU parity: the HIP DPBF through the C-ABI against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

import gpu4gst_code_amd as gst
import oracle

pytestmark = pytest.mark.gpu


def _check_tree(G, groups, r, want_cost, rel=0.0):
    tc = gst.validate_tree(G, groups, r.eu, r.ev, r.root)
    if rel == 0.0:
        assert r.cost == want_cost
        assert tc == want_cost
    else:
        assert abs(r.cost - want_cost) <= rel * max(1.0, abs(want_cost))
        assert abs(tc - want_cost) <= rel * max(1.0, abs(want_cost))


@pytest.mark.parametrize("seed", range(12))
def test_c1_random_connected_bit_exact(solver, seed):
    """V=1000, E=5000, int weights [1,100], g=3 groups of 5-20 members."""
    G = gst.gen_random_connected(1000, 5000, 1, 100, seed=seed)
    groups = gst.gen_groups(G.n, 3, 5, 20, seed=seed + 1)
    dg = solver.upload(G)
    r = solver.solve(dg, groups)
    want = oracle.dpbf_cost(G, groups)
    print(f"seed {seed}: gpu {r.cost} oracle {want} edges {len(r.eu)}", flush=True)
    _check_tree(G, groups, r, want)
    dg.free()


@pytest.mark.parametrize("g", [1, 2, 4, 5, 6, 7, 8])
def test_group_counts(solver, g):
    G = gst.gen_random_connected(600, 2400, 1, 50, seed=100 + g)
    groups = gst.gen_groups(G.n, g, 1, 12, seed=200 + g)
    dg = solver.upload(G)
    r = solver.solve(dg, groups)
    want = oracle.dpbf_cost(G, groups)
    print(f"g {g}: gpu {r.cost} oracle {want}", flush=True)
    _check_tree(G, groups, r, want)
    dg.free()


GOLD = __import__("json").load(open(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "golden.json")))


@pytest.mark.parametrize("pull_permille,chunk", [(-1, 0), (1, 0), (0, 4), (1, 3), (-1, 5)])
def test_push_pull_and_chunking_agree(solver, pull_permille, chunk):
    """Push-only, pull-heavy and tiny work chunks (rows split over many items) give the oracle cost."""
    G = gst.gen_barabasi_albert(5000, 4, 1, 100, seed=31)
    dg = solver.upload(G)
    for q in range(3):
        groups = gst.gen_groups(G.n, 5, 3, 40, seed=40 + q)
        r = solver.solve(dg, groups, pull_permille=pull_permille, chunk_arcs=chunk)
        want = oracle.dpbf_cost(G, groups)
        print(f"pull {pull_permille} chunk {chunk} q{q}: gpu {r.cost} oracle {want}", flush=True)
        _check_tree(G, groups, r, want)
    dg.free()


def test_reference_fixture(solver):
    z = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "gstp_fixture.npz"))
    G = gst.HostGraph(z["row_ptr"], z["col"], z["w"])
    groups = gst.Groups(z["grp_ptr"], z["grp_vtx"])
    dg = solver.upload(G)
    r = solver.solve(dg, groups)
    _check_tree(G, groups, r, GOLD["gstp_fixture"]["cost"])
    dg.free()


def test_ba20k_goldens(solver):
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    dg = solver.upload(G)
    for q, want in GOLD["ba20k"]["costs"].items():
        groups = gst.gen_groups(G.n, 6, 10, 200, seed=1000 + int(q))
        r = solver.solve(dg, groups)
        print(f"ba20k q{q}: gpu {r.cost} golden {want}", flush=True)
        _check_tree(G, groups, r, want)
    dg.free()


def test_float_weights_within_1e6(solver):
    """R-MAT lognormal float32 weights: fp32 DP cost within 1e-6 relative of the double oracle."""
    G = gst.gen_rmat(12, 8, weights="lognormal", seed=21)
    dg = solver.upload(G)
    for q, want in GOLD["rmat_f"]["costs"].items():
        groups = gst.gen_groups(G.n, 5, 5, 50, seed=3000 + int(q))
        r = solver.solve(dg, groups)
        print(f"rmat_f q{q}: gpu {r.cost!r} oracle {want!r}", flush=True)
        assert abs(r.cost - want) <= 1e-6 * want
        tc = gst.validate_tree(G, groups, r.eu, r.ev, r.root)
        assert abs(tc - want) <= 1e-6 * want
    dg.free()


def test_infeasible_and_errors(solver):
    G = gst.HostGraph.from_edges(6, [0, 1, 3, 4], [1, 2, 4, 5], np.array([1, 1, 1, 1], np.int32))
    dg = solver.upload(G)
    with pytest.raises(gst.GstError) as e:
        solver.solve(dg, gst.groups_from_lists([[0], [5]]))
    assert e.value.status == gst._abi.GST_INFEASIBLE
    with pytest.raises(gst.GstError) as e:
        solver.solve(dg, gst.groups_from_lists([[0], []]))
    assert e.value.status == gst._abi.GST_INFEASIBLE
    with pytest.raises(gst.GstError) as e:
        solver.solve(dg, gst.groups_from_lists([[0]] * 13))
    assert e.value.status == gst._abi.GST_TOO_MANY_GROUPS
    with pytest.raises(gst.GstError) as e:
        solver.solve(dg, gst.groups_from_lists([[0], [9]]))
    assert e.value.status == gst._abi.GST_INVALID_ARG
    # the solver stays usable after errors
    r = solver.solve(dg, gst.groups_from_lists([[0], [2]]))
    assert r.cost == 2.0 and gst.validate_tree(G, gst.groups_from_lists([[0], [2]]), r.eu, r.ev, r.root) == 2.0
    # a single vertex covering every group
    r = solver.solve(dg, gst.groups_from_lists([[4], [4, 5], [3, 4]]))
    assert r.cost == 0.0 and len(r.eu) == 0 and r.root == 4
    dg.free()


def test_asymmetric_graph_rejected(solver):
    G = gst.HostGraph(np.array([0, 1, 1], np.int64), np.array([1], np.int32), np.array([3], np.int32))
    with pytest.raises(gst.GstError) as e:
        solver.upload(G)
    assert e.value.status == gst._abi.GST_INVALID_ARG


def test_zero_weight_edges(solver):
    """Zero-weight edges create equal-cost predecessors; the traced structure must still be a tree."""
    for s in range(4):
        G0 = gst.gen_random_connected(400, 1600, 0, 3, seed=70 + s)
        groups = gst.gen_groups(G0.n, 5, 2, 10, seed=80 + s)
        dg = solver.upload(G0)
        r = solver.solve(dg, groups)
        want = oracle.dpbf_cost(G0, groups)
        _check_tree(G0, groups, r, want)
        dg.free()


@pytest.mark.parametrize("slack", [0.0, 0.25])
def test_upper_bound_keeps_exact_cost(solver, slack):
    """A true upper bound (the optimum itself, or 25% above it) prunes states but not the answer."""
    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    dg = solver.upload(G)
    for q in range(3):
        groups = gst.gen_groups(G.n, 6, 10, 200, seed=1000 + q)
        want = GOLD["ba20k"]["costs"][str(q)]
        plain = solver.solve(dg, groups)
        r = solver.solve(dg, groups, ub=want * (1 + slack))
        print(f"ub slack {slack} q{q}: cost {r.cost} relax {r.stats['grow_relax']} vs plain {plain.stats['grow_relax']}",
              flush=True)
        _check_tree(G, groups, r, want)
        assert r.used_ub
        assert r.stats["grow_relax"] < plain.stats["grow_relax"]
    # a bound below the optimum leaves nothing: reported as infeasible under that bound
    groups = gst.gen_groups(G.n, 6, 10, 200, seed=1000)
    with pytest.raises(gst.GstError) as e:
        solver.solve(dg, groups, ub=GOLD["ba20k"]["costs"]["0"] - 1)
    assert e.value.status == gst._abi.GST_INFEASIBLE
    dg.free()


def test_cpp_wrapper_example():
    """The reference-shaped C++ API (include/gst.hpp: DPBF_GPU and DPBF_GPU_T with the reference's
    int* pointer1 / int* real_cost / double* rt, graph_hash_of_mixed_weighted_sum_of_ec,
    this_is_a_feasible_solution_gpu) through the C-ABI; the example exits non-zero unless both
    calls return feasible trees whose int edge sums equal the returned cost, and the MST post-
    processing of the optimal tree's vertex set (graph_hash_of_mixed_weighted_MST_postprocessing_hash_of_
    vertices) keeps that cost."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(__file__)), "gpu4gst_code_amd", "_build", "dpbf_gpu_example")
    for n, m, g, seed in ((1000, 5000, 3, 1), (1000, 5000, 3, 2), (1000, 5000, 3, 3), (1000, 5000, 6, 4)):
        out = subprocess.run([exe, str(n), str(m), str(g), str(seed)], capture_output=True, text=True, timeout=120)
        print(out.stdout, out.stderr, flush=True)
        assert out.returncode == 0
        cost = float(out.stdout.split()[1])
        G = gst.gen_random_connected(n, m, 1, 100, seed=seed)
        groups = gst.gen_groups(G.n, g, 5, 20, seed=seed + 1)
        assert cost == oracle.dpbf_cost(G, groups)


@pytest.mark.parametrize("grow_mode,chunk", [(2, 0), (2, 4), (1, 0), (0, 3)])
def test_work_queue_modes(solver, grow_mode, chunk):
    """Asynchronous work-queue grow (alone, with rows split into many chunk tasks), rounds only,
    and the tuned mix give the oracle cost; g=8 exercises a level wider than one bitmap word."""
    G = gst.gen_barabasi_albert(4000, 4, 1, 60, seed=51)
    dg = solver.upload(G)
    for q, g in enumerate((3, 6, 8)):
        groups = gst.gen_groups(G.n, g, 2, 30, seed=60 + q)
        r = solver.solve(dg, groups, grow_mode=grow_mode, chunk_arcs=chunk)
        want = oracle.dpbf_cost(G, groups)
        print(f"mode {grow_mode} chunk {chunk} g {g}: gpu {r.cost} oracle {want} async {r.stats['async_launches']}",
              flush=True)
        _check_tree(G, groups, r, want)
        rb = solver.solve(dg, groups, ub=want * 1.2, grow_mode=grow_mode, chunk_arcs=chunk)
        _check_tree(G, groups, rb, want)
    dg.free()


def test_work_queue_float_and_zero_weights(solver):
    G = gst.gen_rmat(11, 8, weights="lognormal", seed=23)
    dg = solver.upload(G)
    for q in range(3):
        groups = gst.gen_groups(G.n, 5, 5, 40, seed=90 + q)
        want = oracle.dpbf_cost(G, groups)
        r = solver.solve(dg, groups, grow_mode=2)
        assert abs(r.cost - want) <= 1e-6 * want
    dg.free()
    G0 = gst.gen_random_connected(500, 2000, 0, 2, seed=77)
    dg = solver.upload(G0)
    for q in range(3):
        groups = gst.gen_groups(G0.n, 6, 2, 10, seed=95 + q)
        r = solver.solve(dg, groups, grow_mode=2)
        _check_tree(G0, groups, r, oracle.dpbf_cost(G0, groups))
    dg.free()


@pytest.mark.parametrize("lanes", [1, 8, 32])
def test_push_round_shapes_agree(solver, lanes):
    """Every push-round shape gives the oracle's cost and a valid tree: one item per lane with its
    masks and arcs spread over the wave (lanes = 1) or a lane group per item (8 or 32 lanes);
    unbounded and bounded (U = OPT with the half rule, and the automatic spider bound), one query
    at a time and batched, on int weights with g = 3..8, zero-weight edges and float weights."""
    G = gst.gen_barabasi_albert(5000, 4, 1, 100, seed=31)
    dg = solver.upload(G)
    qs = [gst.gen_groups(G.n, g, 3, 40, seed=60 + g) for g in (3, 5, 6, 8)]
    want = [oracle.dpbf_cost(G, q) for q in qs]
    for q, w in zip(qs, want):
        for kw in ({}, {"ub": w}, {"auto_bound": True}):
            r = solver.solve(dg, q, push_lanes=lanes, pull_permille=-1, **kw)
            print(f"lanes {lanes} g {q.g} {kw}: gpu {r.cost} oracle {w}", flush=True)
            _check_tree(G, q, r, w)
    res, _ = solver.solve_batch(dg, qs + qs, ubs=[None] * len(qs) + want, push_lanes=lanes, slots=5)
    for i, r in enumerate(res):
        _check_tree(G, qs[i % len(qs)], r, want[i % len(qs)])
    dg.free()
    G0 = gst.gen_random_connected(400, 1600, 0, 3, seed=71)
    q0 = gst.gen_groups(G0.n, 5, 2, 10, seed=81)
    dg0 = solver.upload(G0)
    w0 = oracle.dpbf_cost(G0, q0)
    for kw in ({}, {"auto_bound": True}):
        _check_tree(G0, q0, solver.solve(dg0, q0, push_lanes=lanes, pull_permille=-1, **kw), w0)
    dg0.free()
    Gf = gst.gen_rmat(12, 8, weights="lognormal", seed=21)
    qf = gst.gen_groups(Gf.n, 5, 2, 30, seed=52)
    dgf = solver.upload(Gf)
    wf = oracle.dpbf_cost(Gf, qf)
    for kw in ({}, {"auto_bound": True}):
        _check_tree(Gf, qf, solver.solve(dgf, qf, push_lanes=lanes, pull_permille=-1, **kw), wf, rel=1e-6)
    dgf.free()
