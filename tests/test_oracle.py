"""Pins the CPU oracle before it is trusted as the checker.

oracle_cspp (the restated RUN_CSPP_CPU, algorithm/cspp_cpu.hpp:44-125) is
checked against checks that do not share its code path:
  - lexicographic Dijkstra closed forms (bound loose / tight / too small);
  - the resource-expanded DP f[b][v] (a different algorithm);
  - brute-force enumeration of simple paths on tiny graphs;
  - invariants (valid paths, sums, mutually non-dominated Pareto sets,
    monotonicity in the bound) and metamorphic relations (scaling lengths,
    relabelling vertices).
oracle_cspp_astar (label-setting, used for the full-size answer fixtures) is
checked the same way: against oracle_cspp on every grid query and on a
geometric-graph sample, the closed forms, brute force on tiny graphs, and the
committed fixtures' cross-checked queries.
The AR restatements are checked bit-for-bit against golden vectors produced
by the reference's own include/AR.cpp (tests/golden/make_golden.py) and, when
the reference build oracle/_ref exists, against it directly.
"""
import json
import os

import numpy as np
import pytest

from helpers import check_path
from oracle import oracle as O
from vine_amd import workload as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ar_golden.json")


@pytest.fixture(scope="module")
def grid():
    g, q = W.preset("grid64")
    res = [O.cspp(g, q.src[i], q.dst[i], q.bound[i]) for i in range(len(q))]
    return g, q, res


def test_grid_workload_shape(grid):
    g, q, _ = grid
    assert g.n == 4096
    assert g.m == 2 * (8064 - 806)  # 10% of the 8,064 grid edges deleted
    assert len(q) == 100
    assert np.all(q.bound == np.floor(q.mindelay * 12 / 10).astype(np.float32))


def test_closed_forms(grid):
    """P1: bound >= delay of the lexicographic (len, delay) shortest path =>
    answer is that path's (len, delay).  P2: bound < min delay => infeasible.
    P3: bound == min delay => cost is the second key of lexicographic
    Dijkstra on (delay, len)."""
    g, q, res = grid
    n_loose = n_tight = n_inf = 0
    for i in range(len(q)):
        s, t = int(q.src[i]), int(q.dst[i])
        l1, l2 = O.dijkstra_lex(g, s, "len", "delay")
        d1, d2 = O.dijkstra_lex(g, s, "delay", "len")
        r = res[i]
        assert r.feasible == np.isfinite(d1[t])
        if not r.feasible:
            continue
        assert d1[t] == q.mindelay[i]
        if q.bound[i] >= l2[t]:
            n_loose += 1
            assert (r.cost, r.delay) == (l1[t], l2[t])
        # tight bound: only min-delay paths are feasible
        rt = O.cspp(g, s, t, d1[t])
        assert (rt.cost, rt.delay) == (d2[t], d1[t])
        n_tight += 1
        ri = O.cspp(g, s, t, d1[t] - 1)
        assert not ri.feasible
        n_inf += 1
    assert n_tight > 50 and n_inf > 50


def test_resource_dp(grid):
    """P4: the resource-expanded DP gives the same (cost, lex-min delay)."""
    g, q, res = grid
    for i in range(0, len(q), 4):
        f, c, d = O.resource_dp(g, q.src[i], q.dst[i], q.bound[i])
        r = res[i]
        assert f == r.feasible
        if f:
            assert (r.cost, r.delay) == (c, d), i


def _tiny(seed, count):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        n = int(rng.integers(2, 9))
        m = int(rng.integers(1, 3 * n))
        u = rng.integers(0, n, m)
        v = rng.integers(0, n, m)
        keep = u != v
        u, v = u[keep], v[keep]
        order = np.argsort(u, kind="stable")
        u, v = u[order], v[order]
        row = np.zeros(n + 1, np.int32)
        np.add.at(row, u + 1, 1)
        yield W.Graph(np.cumsum(row).astype(np.int32), v.astype(np.int32),
                      rng.integers(0, 10, u.shape[0]).astype(np.float32),
                      rng.integers(1, 10, u.shape[0]).astype(np.float32))


def test_bruteforce_tiny():
    """P5: exhaustive enumeration of simple paths, every prune mode."""
    rng = np.random.default_rng(11)
    checked = 0
    for g in _tiny(3, 200):
        for _ in range(6):
            s, t = int(rng.integers(0, g.n)), int(rng.integers(0, g.n))
            b = float(rng.integers(0, 25))
            f, c, d = O.bruteforce(g, s, t, b)
            for mode in (0, 1, 2):
                r = O.cspp(g, s, t, b, prune_target=mode)
                assert r.feasible == f
                if f:
                    assert (r.cost, r.delay) == (c, d)
                    check_path(g, r.path, s, t, r.cost, r.delay, b)
            checked += f
    assert checked > 300


def test_invariants(grid):
    """P6: paths are valid and sum to the answer; Pareto sets are mutually
    non-dominated and sorted; cost is non-increasing in the bound."""
    g, q, res = grid
    for i in range(0, len(q), 3):
        r = res[i]
        if not r.feasible:
            continue
        check_path(g, r.path, q.src[i], q.dst[i], r.cost, r.delay, q.bound[i])
        p = r.pareto
        assert (p[0, 0], p[0, 1]) == (r.cost, r.delay)
        assert np.all(np.diff(p[:, 0]) > 0) and np.all(np.diff(p[:, 1]) < 0)
        looser = O.cspp(g, q.src[i], q.dst[i], q.bound[i] * 2)
        assert looser.cost <= r.cost


def test_prune_modes_agree(grid):
    g, q, res = grid
    for i in range(0, len(q), 5):
        for mode in (1, 2):
            r = O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=mode)
            assert (r.feasible, r.cost, r.delay) == (res[i].feasible, res[i].cost, res[i].delay)
            assert r.counters["expanded"] <= res[i].counters["expanded"]


def test_metamorphic(grid):
    """P7: doubling every len doubles the cost; relabelling vertices changes
    nothing."""
    g, q, res = grid
    g2 = W.Graph(g.row_ptr, g.dst, g.len * 2, g.delay)
    rng = np.random.default_rng(5)
    perm = rng.permutation(g.n).astype(np.int32)  # old id -> new id
    inv = np.argsort(perm)  # new id -> old id
    row = np.zeros(g.n + 1, np.int32)
    row[1:] = np.cumsum(np.diff(g.row_ptr)[inv])
    idx = np.concatenate([np.arange(g.row_ptr[o], g.row_ptr[o + 1]) for o in inv])
    g3 = W.Graph(row, perm[g.dst[idx]], g.len[idx], g.delay[idx])
    for i in range(0, len(q), 7):
        r2 = O.cspp(g2, q.src[i], q.dst[i], q.bound[i])
        assert r2.feasible == res[i].feasible
        if r2.feasible:
            assert r2.cost == 2 * res[i].cost
        r3 = O.cspp(g3, perm[q.src[i]], perm[q.dst[i]], q.bound[i])
        assert (r3.feasible, r3.cost, r3.delay) == (res[i].feasible, res[i].cost, res[i].delay)


def test_source_is_own_answer():
    g = W.grid(8, 8, 0, seed=1)
    r = O.cspp(g, 5, 5, 0)
    assert r.feasible and r.cost == 0 and r.delay == 0 and list(r.path) == [5]


# --- AR model -----------------------------------------------------------------

@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_ar_burg_golden(golden):
    for case in golden["ar"]:
        c = O.ar_burg(case["series"], 3)
        assert c.tolist() == case["burg"], case["series"][:4]


def test_ar_lsq_golden(golden):
    for case in golden["ar"]:
        c = O.ar_lsq(case["series"], 3)
        np.testing.assert_allclose(c, case["lsq"], rtol=1e-12, atol=1e-12)


def test_ar_against_reference_build(golden):
    if not os.path.exists(O.REF_AR):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    for case in golden["ar"]:
        assert O.ref_ar(case["series"], 3, 0).tolist() == case["burg"]


def test_ar_burg_math():
    """Independent of the reference: an exact AR(1) process x_t = a x_{t-1}
    has Burg first coefficient close to a, and a planted wrong sign fails."""
    a = 0.8
    x = [1.0]
    rng = np.random.default_rng(0)
    for _ in range(4000):
        x.append(a * x[-1] + rng.normal(0, 0.1))
    c = O.ar_burg(x, 1)
    assert abs(c[0] - a) < 0.02
    assert abs(-c[0] - a) > 1.0


def test_label_setting_matches_restated(grid):
    """The label-setting oracle against the restated search on every grid
    query, and the closed forms for tight and too-small bounds (P2, P3)."""
    g, q, res = grid
    for i in range(len(q)):
        r = res[i]
        a = O.cspp_astar(g, q.src[i], q.dst[i], q.bound[i])
        assert (a.feasible, a.cost, a.delay) == (r.feasible, r.cost, r.delay), i
        if not a.feasible:
            continue
        check_path(g, a.path, q.src[i], q.dst[i], a.cost, a.delay, q.bound[i])
        assert a.counters["expanded"] <= r.counters["expanded"]
        d1, d2 = O.dijkstra_lex(g, int(q.src[i]), "delay", "len")
        t = int(q.dst[i])
        at = O.cspp_astar(g, q.src[i], t, d1[t])
        assert (at.feasible, at.cost, at.delay) == (True, d2[t], d1[t])
        assert not O.cspp_astar(g, q.src[i], t, d1[t] - 1).feasible


def test_label_setting_bruteforce_tiny():
    """P5 for the label-setting oracle: parallel arcs, zero lengths, cycles."""
    rng = np.random.default_rng(12)
    checked = 0
    for g in _tiny(4, 200):
        for _ in range(6):
            s, t = int(rng.integers(0, g.n)), int(rng.integers(0, g.n))
            b = float(rng.integers(0, 25))
            f, c, d = O.bruteforce(g, s, t, b)
            a = O.cspp_astar(g, s, t, b)
            assert a.feasible == f
            if f:
                assert (a.cost, a.delay) == (c, d)
                check_path(g, a.path, s, t, a.cost, a.delay, b)
            checked += f
    assert checked > 300


def test_label_setting_geometric_sample():
    """A 20k-vertex 3-NN graph, hop-distance queries with tight bounds: the
    label-setting oracle against the restated search with its bounds."""
    g = W.knn(20000, seed=9)
    q = W.hop_queries(g, 40, hop_lo=5, hop_hi=60, f_permille=(1020, 1400), seed=9)
    for i in range(len(q)):
        r = O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2)
        a = O.cspp_astar(g, q.src[i], q.dst[i], q.bound[i])
        assert (a.feasible, a.cost, a.delay) == (r.feasible, r.cost, r.delay), i
        if a.feasible:
            check_path(g, a.path, q.src[i], q.dst[i], a.cost, a.delay, q.bound[i])


def test_full_fixture_ny_batch():
    """The committed answers of the 4,096-query batched workload: its inputs are
    what the seeded generators produce, the restated search finished and agreed
    on most queries when it was written (when that pass has been run), and a
    sample is re-solved here by both oracles."""
    from helpers import fixture_answers, graph_crc
    fx = fixture_answers("ny_batch")
    g, q = W.preset("ny_batch")
    assert graph_crc(g) == int(fx["graph_crc"])
    assert np.array_equal(fx["src"], q.src) and np.array_equal(fx["dst"], q.dst)
    assert np.array_equal(fx["bound"], q.bound)
    assert fx["feasible"].all()  # hop queries are drawn reachable, bound >= min delay
    if np.all(fx["restated"] == 2):
        # the restated-search pass has not been run on this file: draw from all
        done = np.arange(len(q))
    else:
        done = np.nonzero(fx["restated"] == 1)[0]
        assert done.size > 0.6 * len(q)
    # the cheapest cross-checked queries by the label-setting work, plus a few random ones
    pick = np.concatenate([done[np.argsort(fx["X"][done])[:6]],
                           np.random.default_rng(2).choice(done, 4, replace=False)])
    for i in pick:
        a = O.cspp_astar(g, q.src[i], q.dst[i], q.bound[i])
        assert (a.feasible, np.float32(a.cost), np.float32(a.delay)) == (True, fx["cost"][i], fx["delay"][i])
        assert (a.counters["expanded"], a.counters["relaxed"]) == (fx["X"][i], fx["E"][i])
        assert len(a.path) - 1 == fx["hops"][i]
    for i in pick[:6]:
        r = O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2)
        assert (r.feasible, np.float32(r.cost), np.float32(r.delay)) == (True, fx["cost"][i], fx["delay"][i])
