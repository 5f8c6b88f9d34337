"""Writes the golden vectors under tests/golden/ (run in a container that has the reference mounted).

* gstp_fixture.npz: the reference's only bundled instance, simple_iterative_tests.text (data, not
  code: CSR + group lists), read with gpu4gst_code_amd.gstp.read_gstp.  The reference ships no
  answer for it; its cost here comes from the oracle and is confirmed by the independent g=4
  closed form (two-centre formula over Dijkstra distances).
* golden.json: seed -> optimal cost for the deterministic generators:
    c1      : random connected V=1000 E=5000 int [1,100], g=3 sizes 5..20 (graph seed s, groups s+1);
              oracle cost confirmed by the g=3 spider closed form
    tiny    : V=10..14 random connected, g=1..6, sizes 1..3; confirmed by brute force
    ba20k   : BA V=20000 m=5 int [1,1000], g=6 sizes 10..200 (groups seed 1000+q); oracle only
    rmat_f  : R-MAT scale 12 ef 8 lognormal(0,1) float weights, g=5 sizes 5..50; oracle (double)
Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import gpu4gst_code_amd as gst  # noqa: E402
import oracle  # noqa: E402
from gpu4gst_code_amd.gstp import read_gstp  # noqa: E402

REF_FIXTURE = "/root/reference/simple_iterative_tests.text"


def main():
    out = {}
    if os.path.exists(REF_FIXTURE):
        G, groups, lam = read_gstp(REF_FIXTURE)
        np.savez_compressed(os.path.join(HERE, "gstp_fixture.npz"), row_ptr=G.row_ptr, col=G.col, w=G.w,
                            grp_ptr=groups.grp_ptr, grp_vtx=groups.grp_vtx)
        c = oracle.dpbf_cost(G, groups)
        cf = oracle.closed_form(G, groups)
        assert c == cf, (c, cf)
        out["gstp_fixture"] = {"file": "gstp_fixture.npz", "source": "simple_iterative_tests.text",
                               "n": G.n, "arcs": G.m, "g": groups.g, "cost": c, "confirmed_by": "closed_form_g4"}
    else:
        prev = json.load(open(os.path.join(HERE, "golden.json")))
        out["gstp_fixture"] = prev["gstp_fixture"]

    c1 = {}
    for s in range(200):
        G = gst.gen_random_connected(1000, 5000, 1, 100, seed=s)
        groups = gst.gen_groups(G.n, 3, 5, 20, seed=s + 1)
        c = oracle.dpbf_cost(G, groups)
        assert c == oracle.closed_form(G, groups)
        c1[str(s)] = c
    out["c1"] = {"graph": "random_connected(1000, 5000, 1, 100, seed=s)", "groups": "gen_groups(1000, 3, 5, 20, seed=s+1)",
                 "confirmed_by": "closed_form_g3", "costs": c1}

    tiny = {}
    for s in range(120):
        n = 10 + s % 5
        g = 1 + s % 6
        G = gst.gen_random_connected(n, n + 6, 1, 20, seed=5000 + s)
        groups = gst.gen_groups(n, g, 1, 3, seed=6000 + s)
        c = oracle.dpbf_cost(G, groups)
        assert c == oracle.bruteforce(G, groups)
        tiny[str(s)] = c
    out["tiny"] = {"graph": "random_connected(10 + s%5, n+6, 1, 20, seed=5000+s)",
                   "groups": "gen_groups(n, 1 + s%6, 1, 3, seed=6000+s)", "confirmed_by": "bruteforce", "costs": tiny}

    G = gst.gen_barabasi_albert(20000, 5, 1, 1000, seed=11)
    ba = {}
    for q in range(6):
        groups = gst.gen_groups(G.n, 6, 10, 200, seed=1000 + q)
        ba[str(q)] = oracle.dpbf_cost(G, groups)
    out["ba20k"] = {"graph": "barabasi_albert(20000, 5, 1, 1000, seed=11)", "groups": "gen_groups(20000, 6, 10, 200, seed=1000+q)",
                    "confirmed_by": "oracle", "costs": ba}

    G = gst.gen_rmat(12, 8, weights="lognormal", seed=21)
    rf = {}
    for q in range(6):
        groups = gst.gen_groups(G.n, 5, 5, 50, seed=3000 + q)
        rf[str(q)] = oracle.dpbf_cost(G, groups)
    out["rmat_f"] = {"graph": "rmat(12, 8, lognormal(0,1), seed=21)", "groups": "gen_groups(4096, 5, 5, 50, seed=3000+q)",
                     "confirmed_by": "oracle (double)", "costs": rf}

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote golden.json")


if __name__ == "__main__":
    main()
