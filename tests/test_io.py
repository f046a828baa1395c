"""Graph and query file loaders (the reference's read_graph, include/io.hpp:22-91,
and task::read_query / load_query, include/task.hpp:27-76) on the CPU.

Pinned by hand-built CSR for small files, by round trips through files
written here from generated graphs, and by the reference's shipped inputs
(tests/golden/inputs: tech-routers-rf.mtx.txt and the six rquery_* files, kept
byte-for-byte): their header counts, the weight ranges w in [1, 25], c in
[1, 10], the undirected symmetry, and the rquery invariant
lower <= limit <= upper on all 120 rows.
"""
import os
import subprocess

import numpy as np
import pytest

from vine_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
FACADE = os.path.join(ROOT, "vine_amd", "vine_facade_check")


def write_cola(g, path):
    """COLA text for a CSR graph: one line per vertex with arcs, in CSR order."""
    with open(path, "w") as f:
        f.write("%d\n" % g.n)
        for u in range(g.n):
            lo, hi = g.row_ptr[u], g.row_ptr[u + 1]
            if lo == hi:
                continue
            parts = ["%d %d %d" % (g.dst[p], g.len[p], g.delay[p]) for p in range(lo, hi)]
            f.write("%d: %s -1\n" % (u, " ".join(parts)))


def test_snap_reference_layout_hand_csr(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("3?3?\n0 1 3 4\n1 2 5 6\n0 2 7 8\n")
    g = W.read_graph(p, W.FILE_SNAP)
    # arcs in insertion order: 0->1 1->0 1->2 2->1 0->2 2->0
    assert g.n == 3 and g.m == 6
    assert g.row_ptr.tolist() == [0, 2, 4, 6]
    assert g.dst.tolist() == [1, 2, 0, 2, 1, 0]
    assert g.len.tolist() == [3, 7, 3, 5, 5, 7]
    assert g.delay.tolist() == [4, 8, 4, 6, 6, 8]


def test_snap_matrix_market_hand_csr(tmp_path):
    p = tmp_path / "g.mtx.txt"
    p.write_text("%MatrixMarket matrix coordinate pattern symmetric\n% comment\n4 4 2\n0 3 1 9 2\n1 1 2 4 5\n")
    g = W.read_graph(p, W.FILE_SNAP)
    # 3->1 1->3 1->2 2->1
    assert g.n == 4 and g.row_ptr.tolist() == [0, 0, 2, 3, 4]
    assert g.dst.tolist() == [3, 2, 1, 1]
    assert g.len.tolist() == [9, 4, 4, 9] and g.delay.tolist() == [2, 5, 5, 2]


def test_cola_hand_csr(tmp_path):
    p = tmp_path / "g.graph"
    p.write_text("3\n2: 0 5 1 -1\n0: 1 2 3 2 4 4 -1\n0: 2 9 9 -1\n")
    g = W.read_graph(p, W.FILE_COLA)
    # directed as listed, per vertex in file order: 0->1 0->2 0->2, 2->0
    assert g.n == 3 and g.row_ptr.tolist() == [0, 3, 3, 4]
    assert g.dst.tolist() == [1, 2, 2, 0]
    assert g.len.tolist() == [2, 4, 9, 5] and g.delay.tolist() == [3, 4, 9, 1]


@pytest.mark.parametrize("make", [lambda: W.grid(16, 16, seed=3), lambda: W.rmat(9, 4, seed=2)])
def test_cola_round_trip(tmp_path, make):
    g = make()
    p = tmp_path / "g.graph"
    write_cola(g, p)
    h = W.read_graph(p, W.FILE_COLA)
    assert h.n == g.n
    for a, b in ((h.row_ptr, g.row_ptr), (h.dst, g.dst), (h.len, g.len), (h.delay, g.delay)):
        assert np.array_equal(a, b)


def test_snap_round_trip_of_undirected_generator(tmp_path):
    """Each undirected edge of a generated grid written once as "u v w c" and
    read back gives the same arcs at every vertex."""
    g = W.grid(12, 12, seed=5)
    seen, rows = set(), []
    for u in range(g.n):
        for p in range(g.row_ptr[u], g.row_ptr[u + 1]):
            v = int(g.dst[p])
            key = (min(u, v), max(u, v))
            if key not in seen:
                seen.add(key)
                rows.append((u, v, int(g.len[p]), int(g.delay[p])))
    p = tmp_path / "g.txt"
    with open(p, "w") as f:
        f.write("%d %d\n" % (g.n, len(rows)))
        for r in rows:
            f.write("%d %d %d %d\n" % r)
    h = W.read_graph(p, W.FILE_SNAP)
    assert h.n == g.n and h.m == g.m

    def arcs(x, u):
        lo, hi = x.row_ptr[u], x.row_ptr[u + 1]
        return sorted(zip(x.dst[lo:hi].tolist(), x.len[lo:hi].tolist(), x.delay[lo:hi].tolist()))

    assert all(arcs(h, u) == arcs(g, u) for u in range(g.n))


def test_shipped_tech_routers_rf():
    g = W.read_graph(os.path.join(INPUTS, "tech-routers-rf.mtx.txt"), W.FILE_SNAP)
    assert g.n == 2113 and g.m == 2 * 6632
    assert g.len.min() == 1 and g.len.max() == 25 and g.delay.min() == 1 and g.delay.max() == 10
    # first data row "0 463 0 12 9": arcs 463->0 and 0->463 with (12, 9)
    assert g.dst[g.row_ptr[463]] == 0 and g.len[g.row_ptr[463]] == 12 and g.delay[g.row_ptr[463]] == 9
    assert g.dst[g.row_ptr[0]] == 463
    # undirected: the arc multiset equals its reverse
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    fwd = sorted(zip(src.tolist(), g.dst.tolist(), g.len.tolist(), g.delay.tolist()))
    rev = sorted(zip(g.dst.tolist(), src.tolist(), g.len.tolist(), g.delay.tolist()))
    assert fwd == rev


def test_shipped_rquery_files():
    total = 0
    for name in sorted(os.listdir(INPUTS)):
        if not name.startswith("rquery_"):
            continue
        q = W.read_queries(os.path.join(INPUTS, name), W.FILE_COLA)
        assert len(q) == 20, name
        lower, upper = q.aux[:, 0], q.aux[:, 1]
        assert np.all(lower <= q.bound) and np.all(q.bound <= upper), name
        total += len(q)
    assert total == 120
    q = W.read_queries(os.path.join(INPUTS, "rquery_ny_s1"), W.FILE_COLA)
    assert (q.src[0], q.dst[0], q.bound[0]) == (4968, 5025, 79750)


def test_csv_queries(tmp_path):
    p = tmp_path / "demand.csv"
    p.write_text("id,src,snk,delay,hop,cost\n0,5,7,100,4,9\n1,2,3,50,6,11\n")
    q = W.read_queries(p, W.FILE_SNAP)
    assert q.src.tolist() == [5, 2] and q.dst.tolist() == [7, 3]
    assert q.bound.tolist() == [100, 50] and q.aux.tolist() == [[4, 9], [6, 11]]


@pytest.mark.parametrize("fmt,text,where", [
    (W.FILE_COLA, "3\n0: 1 2 3 -1\n1: 5 1 1 -1\n", ":3:"),          # id out of range
    (W.FILE_COLA, "3\n0: 1 2 -1\n", ":2:"),                          # truncated triple
    (W.FILE_COLA, "3\n0 1 2 3 -1\n", ":2:"),                         # missing ':'
    (W.FILE_SNAP, "3 1\n0 1 -4 2\n", ":2:"),                         # negative weight
    (W.FILE_SNAP, "3 1\n0 1 16777216 2\n", ":2:"),                   # not exact in float32
    (W.FILE_SNAP, "3 1\n0 1 2 2\n1 2 3 3\n", ":3:"),                 # more rows than the header says
    (W.FILE_SNAP, "%MatrixMarket\n3 3 2\n0 0 1 2 2\n", ":4:"),       # fewer rows (found at the end)
    (W.FILE_SNAP, "%MatrixMarket\n3 4 1\n0 0 1 2 2\n", ":2:"),       # not square
])
def test_graph_file_errors(tmp_path, fmt, text, where):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(ValueError) as e:
        W.read_graph(p, fmt)
    assert where in str(e.value), str(e.value)


def test_missing_files_and_bad_query_rows(tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        W.read_graph(tmp_path / "absent", W.FILE_COLA)
    with pytest.raises(ValueError, match="cannot open"):
        W.read_queries(tmp_path / "absent", W.FILE_COLA)
    p = tmp_path / "q.txt"
    p.write_text("1 2 3 4 5\n1 2 3 4\n")
    with pytest.raises(ValueError, match=":2:|:3:"):
        W.read_queries(p, W.FILE_COLA)
    p.write_text("id,src,snk,delay,hop,cost\n0,1,2,-3,4,5\n")
    with pytest.raises(ValueError, match="bound"):
        W.read_queries(p, W.FILE_SNAP)


def test_cpp_facade_reads_files():
    """read_graph and task::read_query of the C++ facade (include/vine/cspp.hpp)
    on the shipped files, through the host-compiled facade program."""
    r = subprocess.run([FACADE, "files", os.path.join(INPUTS, "tech-routers-rf.mtx.txt"), "1",
                        os.path.join(INPUTS, "rquery_ne_s2"), "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split("\n")[:2] == ["graph 2113 13264", "tasks 20"]
