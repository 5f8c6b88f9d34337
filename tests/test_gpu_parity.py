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
