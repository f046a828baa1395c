"""Shared checks for the parity tests."""
import numpy as np


def check_path(g, path, s, t, cost, delay, bound):
    """A returned path must be an s->t vertex sequence along existing arcs such
    that some choice among parallel arcs sums to exactly (cost, delay), with
    delay <= bound.  Sums are float32, added from s towards t as a label
    accumulates them (exact for integer-valued weights in any order)."""
    f32 = np.float32
    path = np.asarray(path)
    assert len(path) >= 1 and path[0] == s and path[-1] == t, (path[:5], s, t)
    sums = {(f32(0), f32(0))}
    for a, b in zip(path[:-1], path[1:]):
        lo, hi = g.row_ptr[a], g.row_ptr[a + 1]
        ks = np.nonzero(g.dst[lo:hi] == b)[0]
        assert ks.size > 0, "no arc %d->%d" % (a, b)
        arcs = {(f32(g.len[lo + k]), f32(g.delay[lo + k])) for k in ks}
        sums = {(x + l, y + d) for (x, y) in sums for (l, d) in arcs if y + d <= f32(bound)}
        assert sums, "path exceeds the bound"
    assert (f32(cost), f32(delay)) in sums, (cost, delay, sorted(sums)[:4])
    assert f32(delay) <= f32(bound)


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


def check_paths_all(g, src, dst, bound, ans):
    """check_path for every query of an answer set, vectorised per path: each
    consecutive vertex pair must be an arc, the float32 sums of the arcs from s
    must equal (cost, delay) and the delay must stay within the bound.  A pair
    joined by parallel arcs goes through check_path."""
    n = g.n
    tails = np.repeat(np.arange(n, dtype=np.int64), np.diff(g.row_ptr))
    keys = tails * n + g.dst.astype(np.int64)
    order = np.argsort(keys, kind="stable")
    skeys = keys[order]
    dup = np.zeros(skeys.shape[0], bool)
    dup[1:] = skeys[1:] == skeys[:-1]
    dup[:-1] |= dup[1:]
    for i in range(len(src)):
        if not ans.feasible[i]:
            assert np.isinf(ans.cost[i]), i
            continue
        p = np.asarray(ans.path_of(i), np.int64)
        assert len(p) >= 1 and p[0] == src[i] and p[-1] == dst[i], (i, p[:4])
        if len(p) == 1:
            assert ans.cost[i] == 0 and ans.delay[i] == 0, i
            continue
        k = p[:-1] * n + p[1:]
        pos = np.minimum(np.searchsorted(skeys, k), skeys.shape[0] - 1)
        assert np.all(skeys[pos] == k), "query %d: a path step is not an arc" % i
        if dup[pos].any():
            check_path(g, p, src[i], dst[i], ans.cost[i], ans.delay[i], bound[i])
            continue
        arcs = order[pos]
        c = np.cumsum(g.len[arcs], dtype=np.float32)
        d = np.cumsum(g.delay[arcs], dtype=np.float32)
        assert c[-1] == ans.cost[i] and d[-1] == ans.delay[i], (i, c[-1], d[-1], ans.cost[i], ans.delay[i])
        assert d[-1] <= np.float32(bound[i]), i


def fixture_answers(name):
    """The oracle's answers for a full-size workload (tests/golden/answers_<name>.npz,
    written by tests/golden/make_full_answers.py)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "answers_%s.npz" % name)
    return dict(np.load(path))


def compare_fixture(fx, ans, idx=None):
    """Every answer bit-exact against the fixture; returns the count compared."""
    idx = np.arange(len(fx["src"])) if idx is None else np.asarray(idx)
    feas = fx["feasible"][idx].astype(bool)
    got = np.asarray(ans.feasible).astype(bool)
    bad = np.nonzero(got != feas)[0]
    assert bad.size == 0, ("feasibility differs", idx[bad[:8]])
    f = np.nonzero(feas)[0]
    bad = f[(ans.cost[f] != fx["cost"][idx][f]) | (ans.delay[f] != fx["delay"][idx][f])]
    assert bad.size == 0, ("cost/delay differ", [(int(idx[j]), float(ans.cost[j]), float(ans.delay[j]),
                                                  float(fx["cost"][idx][j]), float(fx["delay"][idx][j]))
                                                 for j in bad[:8]])
    return len(idx)


def graph_crc(g):
    """CRC-32 of the CSR arrays, as tests/golden/make_full_answers.py records it."""
    import zlib
    c = zlib.crc32(np.ascontiguousarray(g.row_ptr).tobytes())
    for a in (g.dst, g.len, g.delay):
        c = zlib.crc32(np.ascontiguousarray(a).tobytes(), c)
    return c
