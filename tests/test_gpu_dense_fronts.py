"""Large Pareto fronts on the GPU against the oracle.

A layered graph (two vertices per layer, every vertex linked to both of the
next layer's, len and delay anti-correlated) gives each vertex a front of
hundreds of labels, all reached in the same round.  Expanding them sends
dozens of candidates of one wave to the same (query, vertex): they append as
one group whose new chunks form a linked run, and merge folds long sorted
runs into long fronts (growth into new storage, shifts, dominated runs).
Answers, paths and the full Pareto sets at t must match the oracle."""
import numpy as np
import pytest

from helpers import check_path
from oracle import oracle as O
from vine_amd import solver as S
from vine_amd import workload as W

LAYERS, WIDTH = 24, 2


def layered_graph(seed):
    rng = np.random.default_rng(seed)
    n = 2 + LAYERS * WIDTH
    t = n - 1

    def layer(i):
        return [1 + i * WIDTH + j for j in range(WIDTH)]

    arcs = [(0, v) for v in layer(0)]
    for i in range(LAYERS - 1):
        arcs += [(u, v) for u in layer(i) for v in layer(i + 1)]
    arcs += [(u, t) for u in layer(LAYERS - 1)]
    arcs.sort()
    ln = rng.integers(1, 41, len(arcs))
    dl = 41 - ln + rng.integers(0, 3, len(arcs))
    row = np.zeros(n + 1, np.int32)
    for u, _ in arcs:
        row[u + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    dst = np.array([v for _, v in arcs], np.int32)
    return W.from_csr(row, dst, ln.astype(np.float32), dl.astype(np.float32)), t


def queries(g, t):
    src, dst, bnd = [], [], []
    for s_, t_ in ((0, t), (1, t), (3, 2 + 2 * (LAYERS - 3)), (5, 2 + 2 * (LAYERS - 8))):
        md = W.min_delay(g, s_, t_)
        for f in (1.0, 1.1, 1.3, 2.0):
            src.append(s_)
            dst.append(t_)
            bnd.append(np.float32(np.floor(md * f)))
    return np.array(src, np.int32), np.array(dst, np.int32), np.array(bnd, np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("prune", [0, 2])
def test_dense_fronts(gpu, prune):
    g, t = layered_graph(7)
    src, dst, bnd = queries(g, t)
    print("layered graph %d x %d, %d queries, prune %d" % (LAYERS, WIDTH, len(src), prune), flush=True)
    refs = [O.cspp(g, src[i], dst[i], bnd[i]) for i in range(len(src))]
    assert max(len(r.pareto) for r in refs) >= 100  # the fronts this test is about
    # whole-frontier rounds: every label of a vertex is expanded in the same round
    with S.Solver(g, max_batch=len(src), prune=prune, delta_policy=S.DELTA_INF) as s:
        a = s.solve(src, dst, bnd)
        for i, r in enumerate(refs):
            assert bool(a.feasible[i]) == r.feasible, i
            if r.feasible:
                assert (a.cost[i], a.delay[i]) == (np.float32(r.cost), np.float32(r.delay)), i
                check_path(g, a.path_of(i), src[i], dst[i], a.cost[i], a.delay[i], bnd[i])
            if prune == 0:
                assert np.array_equal(s.pareto(i, dst[i]), r.pareto), i
    if prune == 0:
        # candidates of one (query, vertex) in a wave step share chunks
        st = a.stats
        assert 0 < st["cand_chunks"] * 2 < st["candidates"], (st["cand_chunks"], st["candidates"])
