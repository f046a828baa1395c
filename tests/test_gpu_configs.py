"""Full-size workloads on the GPU: every answer checked through size-independent
properties (valid path summing exactly to the cost, delay within the bound,
the same answers under two window policies), and a seeded sample against the
oracle.

C4: R-MAT scale 17 (131,072 vertices, hubs of hundreds of arcs), 500 pairs
from the largest component, bound 1.3x the min delay.
C2: NY-scale 3-NN graph (264,346 vertices), 1,000 random pairs, bounds 1.1x and 1.5x.
C3: NE-scale 3-NN graph (1,524,453 vertices), the first 300 of its 1,000 pairs,
bound 1.2x, learned window against a fixed one, split over three streams.
"""
import numpy as np
import pytest

from helpers import check_path
from oracle import oracle as O
from vine_amd import learn as L
from vine_amd import solver as S
from vine_amd import workload as W


def _solve(g, q, batch, **kw):
    with S.Solver(g, max_batch=batch, prune=S.PRUNE_BOUNDS, rounds_per_sync=16, **kw) as s:
        return s.solve(q.src, q.dst, q.bound)


def _check_all(g, q, a):
    for i in range(len(q)):
        if a.feasible[i]:
            check_path(g, a.path_of(i), q.src[i], q.dst[i], a.cost[i], a.delay[i], q.bound[i])
        else:
            assert np.isinf(a.cost[i])


def _oracle_sample(g, q, a, idx, time_limit):
    checked = 0
    for i in idx:
        try:
            r = O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2, time_limit=time_limit)
        except O.TimeLimit:
            continue
        checked += 1
        assert bool(a.feasible[i]) == r.feasible, i
        if r.feasible:
            assert (a.cost[i], a.delay[i]) == (np.float32(r.cost), np.float32(r.delay)), i
    return checked


@pytest.mark.gpu
def test_rmat17_full(gpu):
    g, q = W.preset("rmat17")
    print("rmat17: n=%d m=%d max out-degree %d, %d queries" % (g.n, g.m, np.diff(g.row_ptr).max(), len(q)),
          flush=True)
    a = _solve(g, q, 500, delta_policy=S.DELTA_FIXED, delta_fixed=20.0)
    b = _solve(g, q, 500, delta_policy=S.DELTA_INF)
    assert np.array_equal(a.cost, b.cost) and np.array_equal(a.delay, b.delay)
    assert np.array_equal(a.feasible, b.feasible)
    assert a.feasible.all()  # pairs come from the largest component, bound >= min delay
    _check_all(g, q, a)
    idx = np.random.default_rng(0).choice(len(q), 24, replace=False)
    assert _oracle_sample(g, q, a, idx, time_limit=5.0) >= 12


@pytest.mark.gpu
def test_ny_random_pairs_full(gpu):
    g, q11, _ = W.preset("ny")
    print("ny: n=%d m=%d, %d queries at 1.1x" % (g.n, g.m, len(q11)), flush=True)
    a = _solve(g, q11, 1000, delta_policy=S.DELTA_FIXED, delta_fixed=12.0)
    b = _solve(g, q11, 1000, delta_policy=S.DELTA_FIXED, delta_fixed=40.0)
    assert np.array_equal(a.cost, b.cost) and np.array_equal(a.delay, b.delay)
    assert np.array_equal(a.feasible, b.feasible)
    _check_all(g, q11, a)
    # the CPU oracle on the shortest queries (long random pairs take it minutes)
    idx = np.argsort(q11.bound)[:8]
    assert _oracle_sample(g, q11, a, idx, time_limit=8.0) >= 6


@pytest.mark.gpu
def test_ny_random_pairs_loose_bounds(gpu):
    g, _, q15 = W.preset("ny")
    print("ny: %d queries at 1.5x" % len(q15), flush=True)
    a = _solve(g, q15, 1000, delta_policy=S.DELTA_FIXED, delta_fixed=12.0)
    b = _solve(g, q15, 1000, delta_policy=S.DELTA_FIXED, delta_fixed=40.0, concurrency=2)
    assert np.array_equal(a.cost, b.cost) and np.array_equal(a.delay, b.delay)
    assert np.array_equal(a.feasible, b.feasible)
    _check_all(g, q15, a)
    idx = np.argsort(q15.bound)[:6]
    assert _oracle_sample(g, q15, a, idx, time_limit=8.0) >= 4


@pytest.mark.gpu
def test_ne_sample_learned_vs_fixed(gpu):
    g, q = W.preset("ne")
    q = q.subset(np.arange(300))
    params = np.load(L.DEFAULT_MODEL)
    print("ne: n=%d m=%d, %d queries at 1.2x" % (g.n, g.m, len(q)), flush=True)
    dg = S.DeviceGraph(g)
    a = _solve(dg, q, 300, delta_policy=S.DELTA_FIXED, delta_fixed=12.0, label_capacity=300_000_000)
    with S.Solver(dg, max_batch=300, prune=S.PRUNE_BOUNDS, rounds_per_sync=16, delta_policy=S.DELTA_MLP,
                  delta_fixed=4.0, concurrency=3, label_capacity=300_000_000) as s:
        s.set_mlp(params)
        b = s.solve(q.src, q.dst, q.bound)
    assert np.array_equal(a.cost, b.cost) and np.array_equal(a.delay, b.delay)
    assert np.array_equal(a.feasible, b.feasible)
    _check_all(g, q, b)
    assert b.stats["expanded"] < a.stats["expanded"]
