"""GSTP and dataset text formats: readers and writers.

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

The dataset triple read by the reference harness (rucgraph/test_gpu.h:115-121):

    <name>.in    first line "V E C"; "input_graph Vertex <v> <node weight> <community>",
                 "input_graph Edge <v1> <v2> <weight>"           (read_input_graph, :157-204)
    <name>.g     one group per line, "g<id>:<member> <member> ...", ids from 1 (read_Group, :243-274)
    <name>.g<k>  one query per line, group vertex ids V + (id - 1) separated by spaces; the last
                 field of a line is not a group (read_inquire, :275-302)

Groups are numbered as the reference numbers them (group id - 1); a query is the list of its group
numbers (the harness turns a field x into group x - V, test_gpu.h:144-147).
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


# ---------------------------------------------------------------------------------------------
# dataset triple (.in / .g / .g<k>)
# ---------------------------------------------------------------------------------------------
def read_dataset_graph(path: str) -> Tuple[HostGraph, list]:
    """<name>.in -> (graph, communities): communities[c] lists the vertices of community c."""
    eu, ev, ew = [], [], []
    comm: dict = {}
    with open(path, "r") as f:
        head = f.readline().split()
        n = int(float(head[0]))
        for line in f:
            p = line.split()
            if len(p) < 4 or p[0] != "input_graph":
                continue
            if p[1] == "Vertex":
                c = int(float(p[4])) if len(p) > 4 else 0
                comm.setdefault(c, []).append(int(p[2]))
            elif p[1] == "Edge":
                eu.append(int(p[2]))
                ev.append(int(p[3]))
                ew.append(float(p[4]))
    w = np.asarray(ew, dtype=np.float64)
    wt = w.astype(np.int32) if (np.all(w == np.round(w)) and (w.size == 0 or w.max() < 2**31)) else w.astype(np.float32)
    a = np.minimum(eu, ev).astype(np.int64)
    b = np.maximum(eu, ev).astype(np.int64)
    key = a * n + b
    order = np.lexsort((wt, key))
    key, wt = key[order], wt[order]
    first = np.ones(key.shape[0], dtype=bool)
    first[1:] = key[1:] != key[:-1]
    key, wt = key[first], wt[first]
    G = HostGraph.from_edges(n, key // n, key % n, wt)
    communities = [sorted(comm[c]) for c in sorted(comm)]
    return G, communities


def read_dataset_groups(path: str) -> dict:
    """<name>.g -> {group number (id - 1): sorted member list}."""
    out = {}
    with open(path, "r") as f:
        for line in f:
            if ":" not in line:
                continue
            head, rest = line.split(":", 1)
            gid = int(float(head.strip()[1:])) - 1
            out[gid] = sorted(set(int(x) for x in rest.split()))
    return out


def read_dataset_queries(path: str, n: int) -> list:
    """<name>.g<k> -> list of queries, each the list of its group numbers."""
    out = []
    with open(path, "r") as f:
        for line in f:
            p = line.rstrip("\r\n").split(" ")   # the reference writes a space after the last group
            if not line.strip():
                continue
            out.append([int(x) - n for x in p[:-1] if x != ""])
    return out


def dataset_query(groups_table: dict, query) -> Groups:
    """The Groups of one query: group i of the solve is the query's i-th group number."""
    lists = [groups_table[q] for q in query]
    ptr = np.zeros(len(lists) + 1, dtype=np.int32)
    for i, l in enumerate(lists):
        ptr[i + 1] = ptr[i] + len(l)
    vtx = np.asarray([v for l in lists for v in l], dtype=np.int32)
    return Groups(ptr, vtx)


def write_dataset(prefix: str, G: HostGraph, groups_table: dict, queries, suffix: str = ".g7",
                  communities=None) -> None:
    """Write <prefix>.in, <prefix>.g and <prefix><suffix> in the reference harness's formats."""
    n = G.n
    comm_of = np.zeros(n, dtype=np.int64)
    for c, vs in enumerate(communities or []):
        comm_of[np.asarray(vs, dtype=np.int64)] = c
    with open(prefix + ".in", "w") as f:
        f.write(f"{n} {G.m // 2} {max(1, len(communities or []))}\n")
        for v in range(n):
            f.write(f"input_graph Vertex {v} 0 {int(comm_of[v])}\n")
        for v in range(n):
            for e in range(G.row_ptr[v], G.row_ptr[v + 1]):
                u = int(G.col[e])
                if v < u:
                    f.write(f"input_graph Edge {v} {u} {G.w[e]}\n")
    with open(prefix + ".g", "w") as f:
        for gid in sorted(groups_table):
            f.write(f"g{gid + 1}:" + " ".join(str(int(m)) for m in groups_table[gid]) + "\n")
    with open(prefix + suffix, "w") as f:
        for q in queries:
            f.write(" ".join(str(n + int(x)) for x in q) + " \n")
