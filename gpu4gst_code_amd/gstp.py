"""GSTP text format reader/writer.

Reads the instance format of graph_hash_of_mixed_weighted_read_for_GSTP
(rucgraph/graph_hash_of_mixed_weighted_read_for_GSTP.h:9-73):

    input_graph Vertex <v> <node weight>
    input_graph Edge <v1> <v2> <weight>
    group_graph Edge <member> <group vertex> <1>
    group_vertices <group vertex>
    lambda <value>

Input vertices are 0..V-1; group vertices are ids >= V whose group_graph edges name the members.
Groups are numbered in increasing group-vertex id (the reference iterates an unordered_set, whose
order is unspecified; the optimal cost does not depend on it).  Unlike the reference reader, which
truncates edge weights to int (:39), weights stay exact: int32 when every weight is integral,
float32 otherwise.  The writer follows graph_hash_of_mixed_weighted_save_for_GSTP
(rucgraph/graph_hash_of_mixed_weighted_save_for_GSTP.h:5-101) for the fields above.
"""
from __future__ import annotations

from typing import Tuple

import numpy as np

from . import Groups, HostGraph


def read_gstp(path: str) -> Tuple[HostGraph, Groups, int]:
    eu, ev, ew = [], [], []
    members = {}
    group_ids = []
    max_v = -1
    lam = 1
    with open(path, "r") as f:
        for line in f:
            p = line.split()
            if len(p) < 2:
                continue
            if p[0] == "input_graph" and p[1] == "Vertex":
                max_v = max(max_v, int(p[2]))
            elif p[0] == "input_graph" and p[1] == "Edge":
                a, b = int(p[2]), int(p[3])
                eu.append(a)
                ev.append(b)
                ew.append(float(p[4]))
                max_v = max(max_v, a, b)
            elif p[0] == "group_graph" and p[1] == "Edge":
                a, b = int(p[2]), int(p[3])
                # one endpoint is the group vertex (the larger id)
                m, gv = (a, b) if a < b else (b, a)
                members.setdefault(gv, []).append(m)
            elif p[0] == "group_vertices":
                group_ids.append(int(p[1]))
            elif p[0] == "lambda":
                lam = int(float(p[1]))
    n = max_v + 1
    w = np.asarray(ew, dtype=np.float64)
    if np.all(w == np.round(w)) and (w.size == 0 or w.max() < 2**31):
        wt = w.astype(np.int32)
    else:
        wt = w.astype(np.float32)
    # undirected edge list may hold both directions or repeats: keep one copy, lightest weight
    a = np.minimum(eu, ev).astype(np.int64)
    b = np.maximum(eu, ev).astype(np.int64)
    key = a * n + b
    order = np.lexsort((wt, key))
    key, wt = key[order], wt[order]
    first = np.ones(key.shape[0], dtype=bool)
    first[1:] = key[1:] != key[:-1]
    key, wt = key[first], wt[first]
    G = HostGraph.from_edges(n, key // n, key % n, wt)
    lists = [sorted(set(members.get(gv, []))) for gv in sorted(group_ids)]
    ptr = np.zeros(len(lists) + 1, dtype=np.int32)
    for i, l in enumerate(lists):
        ptr[i + 1] = ptr[i] + len(l)
    vtx = np.asarray([v for l in lists for v in l], dtype=np.int32)
    return G, Groups(ptr, vtx), lam


def write_gstp(path: str, G: HostGraph, groups: Groups, lam: int = 1) -> None:
    n = G.n
    with open(path, "w") as f:
        f.write("SECTION Comments\nName \"%s\"\nCreator \"gpu4gst_code_amd.gstp\"\nEND\n\n" % path)
        f.write(f"lambda {lam}\n\n")
        f.write(f"input_graph |V|={n} |E|={G.m // 2}\n")
        for v in range(n):
            f.write(f"input_graph Vertex {v} 0.0000000000\n")
        f.write("\n")
        for v in range(n):
            for e in range(G.row_ptr[v], G.row_ptr[v + 1]):
                u = int(G.col[e])
                if v < u:
                    f.write(f"input_graph Edge {v} {u} {G.w[e]}\n")
        f.write("\n")
        for i in range(groups.g):
            for m in groups.members(i):
                f.write(f"group_graph Edge {int(m)} {n + i} 1\n")
        f.write(f"\n|T|={groups.g}\n")
        for i in range(groups.g):
            f.write(f"group_vertices {n + i}\n")
        f.write("\nEOF\n")
