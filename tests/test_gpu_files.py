"""Graphs and queries read from files, solved on the GPU, against the oracle.

The shipped tech-routers-rf graph (tests/golden/inputs, N = 2,113, read by
the MatrixMarket branch of the SNAP loader) with seeded queries at three
bound factors, every answer checked bit-exact; and both file formats end to
end through the C++ facade's read_graph / task::read_query: a COLA graph and
query file written from a generated graph, and a CSV demand file on the
shipped graph.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import check_path
from oracle import oracle as O
from vine_amd import solver as S
from vine_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
FACADE = os.path.join(ROOT, "vine_amd", "vine_facade_check")


@pytest.fixture(scope="module")
def routers():
    return W.read_graph(os.path.join(INPUTS, "tech-routers-rf.mtx.txt"), W.FILE_SNAP)


def _oracle_lines(g, src, dst, bound):
    out = []
    for s, t, b in zip(src, dst, bound):
        r = O.cspp(g, int(s), int(t), float(b), prune_target=2)
        out.append("%d %.0f %.0f" % (1 if r.feasible else 0, r.cost if r.feasible else np.inf,
                                     r.delay if r.feasible else np.inf))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("factor", [(10, 10), (12, 10), (20, 10)])
def test_routers_parity(gpu, routers, factor):
    g = routers
    q = W.random_queries(g, 300, factor, seed=factor[0])
    print("tech-routers-rf: %d queries at %d/%d x min delay" % (len(q), factor[0], factor[1]), flush=True)
    with S.Solver(g, max_batch=300, prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_FIXED, delta_fixed=8.0) as s:
        a = s.solve(q.src, q.dst, q.bound)
    for i in range(len(q)):
        r = O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2)
        assert bool(a.feasible[i]) == r.feasible, i
        if r.feasible:
            assert (a.cost[i], a.delay[i]) == (np.float32(r.cost), np.float32(r.delay)), i
            check_path(g, a.path_of(i), q.src[i], q.dst[i], a.cost[i], a.delay[i], q.bound[i])


@pytest.mark.gpu
def test_facade_cola_files(gpu, tmp_path):
    g = W.knn(3000, seed=7)
    q = W.random_queries(g, 40, (13, 10), seed=7)
    gp, qp = tmp_path / "knn.graph", tmp_path / "rquery"
    with open(gp, "w") as f:
        f.write("%d\n" % g.n)
        for u in range(g.n):
            lo, hi = g.row_ptr[u], g.row_ptr[u + 1]
            f.write("%d: %s -1\n" % (u, " ".join("%d %d %d" % (g.dst[p], g.len[p], g.delay[p])
                                                 for p in range(lo, hi))))
    with open(qp, "w") as f:
        for i in range(len(q)):
            f.write("%d %d %d %d %d\n" % (q.src[i], q.dst[i], max(q.mindelay[i], 0), int(q.bound[i]),
                                          int(q.bound[i])))
    print("facade COLA files: n=%d, %d queries" % (g.n, len(q)), flush=True)
    r = subprocess.run([FACADE, "files", str(gp), "2", str(qp), "2", "solve"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[:2] == ["graph %d %d" % (g.n, g.m), "tasks %d" % len(q)]
    assert lines[2:] == _oracle_lines(g, q.src, q.dst, q.bound)


@pytest.mark.gpu
def test_facade_csv_demand_on_shipped_graph(gpu, routers, tmp_path):
    g = routers
    q = W.random_queries(g, 30, (11, 10), seed=3)
    qp = tmp_path / "demand.csv"
    with open(qp, "w") as f:
        f.write("id,src,snk,delay,hop,cost\n")
        for i in range(len(q)):
            f.write("%d,%d,%d,%d,0,0\n" % (i, q.src[i], q.dst[i], int(q.bound[i])))
    print("facade CSV demand on tech-routers-rf: %d queries" % len(q), flush=True)
    r = subprocess.run([FACADE, "files", os.path.join(INPUTS, "tech-routers-rf.mtx.txt"), "1", str(qp), "1",
                        "solve"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[:2] == ["graph 2113 13264", "tasks %d" % len(q)]
    assert lines[2:] == _oracle_lines(g, q.src, q.dst, q.bound)
