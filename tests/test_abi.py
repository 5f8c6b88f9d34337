"""CPU: the C-ABI library loads, exports every declared symbol, and its host logic behaves."""
import os
import re

import numpy as np
import pytest

import gpu4gst_code_amd as gst
from gpu4gst_code_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gst_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names <= set(_abi.SIGNATURES), names - set(_abi.SIGNATURES)


def test_status_strings():
    lib = _abi.load()
    for code, name in _abi.STATUS_NAMES.items():
        assert lib.gst_status_str(code).decode() == name


def test_no_device_fails_loudly():
    """Without a GPU the solver refuses to run (there is no CPU fallback)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(gst.GstError) as e:
        gst.Solver(0)
    assert e.value.status == _abi.GST_NO_DEVICE


def _digest(G):
    return hash((G.row_ptr.tobytes(), G.col.tobytes(), G.w.tobytes()))


def test_generators_deterministic_and_shaped():
    a = gst.gen_random_connected(1000, 5000, 1, 100, seed=5)
    b = gst.gen_random_connected(1000, 5000, 1, 100, seed=5)
    c = gst.gen_random_connected(1000, 5000, 1, 100, seed=6)
    assert _digest(a) == _digest(b) != _digest(c)
    assert a.n == 1000 and a.m == 10000
    assert a.w.min() >= 1 and a.w.max() <= 100
    ba = gst.gen_barabasi_albert(2000, 5, 1, 1000, seed=1)
    assert ba.m == 2 * (10 + (2000 - 5) * 5)
    r = gst.gen_rmat(10, 8, weights="lognormal", seed=2)
    assert r.n == 1024 and r.w.dtype == np.float32 and (r.w > 0).all()
    g1 = gst.gen_groups(1000, 6, 10, 200, seed=9)
    g2 = gst.gen_groups(1000, 6, 10, 200, seed=9)
    assert np.array_equal(g1.grp_vtx, g2.grp_vtx) and g1.g == 6
    for i in range(6):
        m = g1.members(i)
        assert 10 <= len(m) <= 200 and len(set(m.tolist())) == len(m)


def _symmetric(G):
    src = np.repeat(np.arange(G.n), np.diff(G.row_ptr))
    fwd = set(zip(src.tolist(), G.col.tolist(), G.w.tolist()))
    bwd = set(zip(G.col.tolist(), src.tolist(), G.w.tolist()))
    return fwd == bwd and not np.any(src == G.col)


def test_generated_graphs_are_symmetric_simple():
    for G in (gst.gen_random_connected(300, 900, 1, 9, seed=1), gst.gen_barabasi_albert(500, 3, 1, 9, seed=2),
              gst.gen_rmat(9, 8, seed=3)):
        assert _symmetric(G)
        for v in range(G.n):
            row = G.col[G.row_ptr[v]:G.row_ptr[v + 1]]
            assert np.all(np.diff(row) > 0)


def test_validate_tree_rejects_bad_trees():
    G = gst.HostGraph.from_edges(5, [0, 1, 2, 3, 0], [1, 2, 3, 4, 4], np.array([1, 2, 3, 4, 5], np.int32))
    groups = gst.groups_from_lists([[0], [3]])
    assert gst.validate_tree(G, groups, [0, 1, 2], [1, 2, 3]) == 6.0
    assert gst.validate_tree(G, groups, [0, 4], [4, 3]) == 9.0
    with pytest.raises(gst.GstError):  # cycle
        gst.validate_tree(G, groups, [0, 1, 2, 3, 0], [1, 2, 3, 4, 4])
    with pytest.raises(gst.GstError):  # disconnected
        gst.validate_tree(G, groups, [0, 2], [1, 3])
    with pytest.raises(gst.GstError):  # not an edge
        gst.validate_tree(G, groups, [0], [3])
    with pytest.raises(gst.GstError):  # group 1 untouched
        gst.validate_tree(G, groups, [0], [1])
    single = gst.groups_from_lists([[2], [2, 4]])
    assert gst.validate_tree(G, single, [], [], root=2) == 0.0


def test_gstp_roundtrip(tmp_path):
    from gpu4gst_code_amd.gstp import read_gstp, write_gstp
    G = gst.gen_random_connected(50, 120, 1, 9, seed=4)
    groups = gst.gen_groups(50, 3, 2, 5, seed=5)
    p = tmp_path / "x.stp"
    write_gstp(str(p), G, groups)
    G2, groups2, lam = read_gstp(str(p))
    assert np.array_equal(G.row_ptr, G2.row_ptr) and np.array_equal(G.col, G2.col) and np.array_equal(G.w, G2.w)
    for i in range(3):
        assert sorted(groups.members(i).tolist()) == groups2.members(i).tolist()


def test_dataset_triple_round_trip(tmp_path):
    """The reference harness's .in / .g / .g7 files (rucgraph/test_gpu.h:115-121): write, read back,
    and the queries' optima from the files equal those of the original groups."""
    import oracle
    from gpu4gst_code_amd import gstp
    G = gst.gen_random_connected(300, 1200, 1, 50, seed=9)
    rng = np.random.default_rng(3)
    table = {i: sorted(set(rng.choice(G.n, size=int(rng.integers(2, 9)), replace=False).tolist())) for i in range(12)}
    queries = [sorted(rng.choice(12, size=4, replace=False).tolist()) for _ in range(5)]
    prefix = str(tmp_path / "toy")
    gstp.write_dataset(prefix, G, table, queries, communities=[list(range(150)), list(range(150, 300))])
    G2, comm = gstp.read_dataset_graph(prefix + ".in")
    assert G2.n == G.n and G2.m == G.m
    assert np.array_equal(G2.row_ptr, G.row_ptr) and np.array_equal(G2.col, G.col) and np.array_equal(G2.w, G.w)
    assert [len(c) for c in comm] == [150, 150]
    t2 = gstp.read_dataset_groups(prefix + ".g")
    assert t2 == table
    q2 = gstp.read_dataset_queries(prefix + ".g7", G2.n)
    assert q2 == queries
    for q in q2:
        a = gstp.dataset_query(t2, q)
        b = gst.groups_from_lists([table[i] for i in q])
        assert oracle.dpbf_cost(G2, a) == oracle.dpbf_cost(G, b)


def test_dataset_reader_reference_lines(tmp_path):
    """Lines as the reference writes them: header "V E C", community field, g<id>: rows, query
    lines ending in a field that is not a group."""
    from gpu4gst_code_amd import gstp
    (tmp_path / "d.in").write_text("4 3 2\ninput_graph Vertex 0 0 0\ninput_graph Vertex 1 0 0\n"
                                   "input_graph Vertex 2 0 1\ninput_graph Vertex 3 0 1\n"
                                   "input_graph Edge 0 1 5\ninput_graph Edge 1 2 7\ninput_graph Edge 2 3 2\n")
    (tmp_path / "d.g").write_text("g1:0 3\ng2:2\ng3:1 2\n")
    (tmp_path / "d.g7").write_text("4 5 6 \n6 4 \n")
    G, comm = gstp.read_dataset_graph(str(tmp_path / "d.in"))
    assert G.n == 4 and G.m == 6 and comm == [[0, 1], [2, 3]]
    t = gstp.read_dataset_groups(str(tmp_path / "d.g"))
    assert t == {0: [0, 3], 1: [2], 2: [1, 2]}
    q = gstp.read_dataset_queries(str(tmp_path / "d.g7"), G.n)
    assert q == [[0, 1, 2], [2, 0]]
    grp = gstp.dataset_query(t, q[0])
    assert grp.g == 3 and list(grp.members(0)) == [0, 3]


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors in gpu4gst_code_amd/_abi.py have the sizes and field offsets the C compiler
    gives the structs of include/gst.h (a field added to one side only fails here)."""
    import ctypes as C
    import shutil
    import subprocess
    from gpu4gst_code_amd import _abi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"gst_csr": _abi.gst_csr, "gst_groups": _abi.gst_groups, "gst_opts": _abi.gst_opts,
               "gst_stats": _abi.gst_stats, "gst_result": _abi.gst_result, "gst_learned_stats": _abi.gst_learned_stats,
               "gst_batch_opts": _abi.gst_batch_opts, "gst_batch_stats": _abi.gst_batch_stats,
               "gst_tune_report": _abi.gst_tune_report, "gst_host_csr": _abi.gst_host_csr}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gst.h"', 'int main(void) {']
    for name, cls in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    lines.append('return 0; }')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for name, cls in structs.items():
        assert int(got[f"{name} size"]) == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, (name, f)
