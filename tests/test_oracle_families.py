"""Pins of the label-family oracle (oracle/families_oracle.c) on the CPU.

- Every family against a Python enumeration of all simple s->t paths on tiny
  random graphs, with each path's label computed from its arcs as the
  reference's classes define it (include/label.hpp:46-169, 248-461: sums, hop
  count, longest arc, products from 1, bottleneck from 21474836), the
  constraint checked on every prefix (include/constraint.hpp:21-80, 106-186),
  and the non-dominated set taken by brute force: t's whole Pareto set must
  match, in the answer order.
- The C brute force (oracle_family_bruteforce) against the same sets.
- QoS against Dijkstra (scipy) on the arcs whose delay meets the bound.
- The hop family with loose limits, and the 2-OSP / CSP families, against
  the c_label oracle and Dijkstra.
The social family is compared only with delay_L >= 1: below 1 its constraint
runs against its dominance order (a dominating label has the larger delay),
so a label search's result depends on the order labels meet, in the
reference as here.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

from oracle import oracle as O
from vine_amd import workload as W

F32 = np.float32
CSP, HOP, HOPMAX, OSP2, SOCIAL, QOS = 1, 2, 3, 4, 5, 6


def tiny_graph(rng, n, m, social=False):
    arcs = set()
    while len(arcs) < m:
        u, v = rng.integers(0, n, 2)
        if u != v:
            arcs.add((int(u), int(v)))
    arcs = sorted(arcs)
    if social:
        ln = rng.integers(1, 9, len(arcs)) / 8.0
        dl = rng.integers(1, 9, len(arcs)) / 8.0
    else:
        ln = rng.integers(1, 10, len(arcs)).astype(float)
        dl = rng.integers(1, 10, len(arcs)).astype(float)
    row = np.zeros(n + 1, np.int32)
    for u, _ in arcs:
        row[u + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    return W.from_csr(row, np.array([v for _, v in arcs], np.int32), ln.astype(F32), dl.astype(F32))


def path_label(fam, g, arcs):
    """Labels of every prefix of a path (list of arc positions), each computed
    from the arcs as the family defines it."""
    d = F32(1) if fam == SOCIAL else F32(0)
    dl = F32(1) if fam == SOCIAL else (F32(21474836) if fam == QOS else F32(0))
    hop, mc = 0, F32(0)
    out = []
    for p in arcs:
        if fam == SOCIAL:
            d, dl = F32(d * g.len[p]), F32(dl * g.delay[p])
        elif fam == QOS:
            d, dl = F32(d + g.len[p]), min(dl, g.delay[p])
        else:
            d, dl = F32(d + g.len[p]), F32(dl + g.delay[p])
        hop += 1
        mc = max(mc, g.len[p])
        lab = [d, dl, F32(hop) if fam in (HOP, HOPMAX) else F32(0), mc if fam == HOPMAX else F32(0)]
        out.append(lab)
    return out


def passes(fam, lab, lim):
    if fam == QOS:
        return lab[1] >= lim[0]
    ok = lab[1] <= lim[0]
    if fam in (HOP, HOPMAX):
        ok = ok and lab[2] <= int(lim[1])
    if fam == HOPMAX:
        ok = ok and lab[3] <= lim[2]
    return ok


def dominates(fam, a, b):
    if fam == SOCIAL:
        return a[0] >= b[0] and a[1] >= b[1]
    if fam == QOS:
        return a[0] <= b[0] and a[1] >= b[1]
    k = {HOP: 3, HOPMAX: 4}.get(fam, 2)
    return all(a[i] <= b[i] for i in range(k))


def answer_key(fam, lab):
    if fam == SOCIAL:
        return (-lab[0], -lab[1])
    if fam == QOS:
        return (lab[0], -lab[1])
    return tuple(lab[:{HOP: 3, HOPMAX: 4}.get(fam, 2)])


def python_front(fam, g, s, t, lim):
    labels = []
    if s == t:
        src = [F32(1), F32(1), F32(0), F32(0)] if fam == SOCIAL else \
              [F32(0), F32(21474836) if fam == QOS else F32(0), F32(0), F32(0)]
        labels.append(src)

    def dfs(u, on, arcs):
        for p in range(g.row_ptr[u], g.row_ptr[u + 1]):
            v = int(g.dst[p])
            if v in on:
                continue
            labs = path_label(fam, g, arcs + [p])
            if not passes(fam, labs[-1], lim):
                continue
            if v == t:
                labels.append(labs[-1])
            else:
                dfs(v, on | {v}, arcs + [p])

    if s != t:
        dfs(s, {s}, [])
    front = []
    for i, a in enumerate(labels):
        if any(dominates(fam, b, a) and (not dominates(fam, a, b) or j < i) for j, b in enumerate(labels) if j != i):
            continue
        front.append(a)
    front.sort(key=lambda lab: answer_key(fam, lab))
    return np.array(front, np.float32).reshape(-1, 4)


CASES = [(fam, seed) for fam in (CSP, HOP, HOPMAX, OSP2, SOCIAL, QOS) for seed in range(6)]


@pytest.mark.parametrize("fam,seed", CASES)
def test_family_oracle_matches_enumeration(fam, seed):
    rng = np.random.default_rng(100 * fam + seed)
    g = tiny_graph(rng, 8, 26, social=fam == SOCIAL)
    for _ in range(6):
        s, t = (int(x) for x in rng.integers(0, g.n, 2))
        if fam == SOCIAL:
            lim = (float(rng.choice([1.0, 2.0])), 100, 10000)
        elif fam == QOS:
            lim = (float(rng.integers(1, 8)), 100, 10000)
        else:
            lim = (float(rng.integers(5, 40)), int(rng.integers(1, 6)), float(rng.integers(3, 10)))
        exp = python_front(fam, g, s, t, lim)
        r = O.family(g, fam, s, t, lim)
        assert np.array_equal(r.front, exp), (fam, s, t, lim, r.front, exp)
        assert r.feasible == (len(exp) > 0)
        if r.feasible:
            assert np.array_equal(r.best, exp[0])
        bf = O.family_bruteforce(g, fam, s, t, lim)
        assert np.array_equal(bf, exp)


def test_qos_against_dijkstra_on_wide_arcs():
    g = W.grid(16, 16, delay_range=(1, 20), seed=4)
    rng = np.random.default_rng(4)
    for _ in range(20):
        s, t = (int(x) for x in rng.integers(0, g.n, 2))
        L = float(rng.integers(1, 12))
        keep = g.delay >= L
        src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
        A = sp.csr_matrix((g.len[keep].astype(np.float64), (src[keep], g.dst[keep])), shape=(g.n, g.n))
        dist = dijkstra(A, indices=s)[t]
        r = O.family(g, QOS, s, t, (L, 100, 10000))
        if np.isinf(dist):
            assert not r.feasible
        else:
            assert r.feasible and r.best[0] == dist
            # the answer's bottleneck is the best among the shortest paths: at least L
            assert r.best[1] >= L


def test_csp_families_agree_with_c_label_oracle():
    g, q = W.preset("grid64")
    for i in range(0, 100, 7):
        ref = O.cspp(g, q.src[i], q.dst[i], q.bound[i])
        for fam in (CSP, OSP2):
            r = O.family(g, fam, q.src[i], q.dst[i], (q.bound[i], 0, 0))
            assert r.feasible == ref.feasible
            if ref.feasible:
                assert (r.best[0], r.best[1]) == (F32(ref.cost), F32(ref.delay))
                assert np.array_equal(r.front[:, :2], ref.pareto)


def test_hop_family_with_loose_limits_is_dijkstra():
    g = W.grid(12, 12, seed=9)
    rng = np.random.default_rng(9)
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    A = sp.csr_matrix((g.len.astype(np.float64), (src, g.dst)), shape=(g.n, g.n))
    for _ in range(10):
        s, t = (int(x) for x in rng.integers(0, g.n, 2))
        r = O.family(g, HOP, s, t, (1e7, 10 ** 6, 0))
        assert r.feasible and r.best[0] == dijkstra(A, indices=s)[t]
        # a hop limit below the hop distance makes the query infeasible
        hops = dijkstra(sp.csr_matrix((np.ones(g.m), (src, g.dst)), shape=(g.n, g.n)), indices=s)[t]
        if hops > 0:
            assert not O.family(g, HOP, s, t, (1e7, int(hops) - 1, 0)).feasible
            assert O.family(g, HOP, s, t, (1e7, int(hops), 0)).feasible
