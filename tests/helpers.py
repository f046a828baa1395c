"""Shared checks for the parity tests."""
import numpy as np


def check_path(g, path, s, t, cost, delay, bound):
    """A returned path must be an s->t vertex sequence along existing arcs such
    that some choice among parallel arcs sums to exactly (cost, delay), with
    delay <= bound.  Sums of integer-valued weights are exact."""
    path = np.asarray(path)
    assert len(path) >= 1 and path[0] == s and path[-1] == t, (path[:5], s, t)
    sums = {(0.0, 0.0)}
    for a, b in zip(path[:-1], path[1:]):
        lo, hi = g.row_ptr[a], g.row_ptr[a + 1]
        ks = np.nonzero(g.dst[lo:hi] == b)[0]
        assert ks.size > 0, "no arc %d->%d" % (a, b)
        arcs = {(float(g.len[lo + k]), float(g.delay[lo + k])) for k in ks}
        sums = {(x + l, y + d) for (x, y) in sums for (l, d) in arcs if y + d <= bound}
        assert sums, "path exceeds the bound"
    assert (float(cost), float(delay)) in sums, (cost, delay, sorted(sums)[:4])
    assert float(delay) <= float(bound)


def facade_grid():
    """The 4x4 grid that tests/cpp/facade_main.cpp builds (undirected edges,
    adjacency insertion order)."""
    from vine_amd import workload as W
    arcs = []
    for rr in range(4):
        for cc in range(4):
            i = rr * 4 + cc
            if cc + 1 < 4:
                w = (1.0 + (i * 7) % 5, 1.0 + (i * 3) % 4)
                arcs += [(i, i + 1, w), (i + 1, i, w)]
            if rr + 1 < 4:
                w = (2.0 + (i * 5) % 3, 1.0 + (i * 11) % 3)
                arcs += [(i, i + 4, w), (i + 4, i, w)]
    adj = [[] for _ in range(16)]
    for u, v, w in arcs:
        adj[u].append((v, w))
    row = np.zeros(17, np.int32)
    dst, ln, dl = [], [], []
    for u in range(16):
        for v, w in adj[u]:
            dst.append(v)
            ln.append(w[0])
            dl.append(w[1])
        row[u + 1] = len(dst)
    return W.Graph(row, np.array(dst, np.int32), np.array(ln, np.float32), np.array(dl, np.float32))
