"""The batched workload at full size: 4,096 queries on the NY-scale 3-NN graph
(264,346 vertices), hop distance log-uniform in [10, 2000], bound
U[1.05, 2.0] x the min delay, learned window, three streams.  The oracle
takes minutes on the long queries, so every answer is pinned by relations
that hold at any size:
- every returned path is an arc sequence s -> t whose lengths sum exactly to
  the cost and whose delays sum to the delay, within the bound;
- doubling every length doubles every cost and leaves delays and
  feasibility unchanged (integer sums below 2^24 are exact);
- renumbering the vertices (a seeded permutation, each vertex's arcs listed
  in reverse) and solving with a fixed window instead of the learned one
  gives the same (cost, delay) for every query;
- a looser bound never raises a cost and never loses a feasible query."""
import numpy as np
import pytest

from helpers import check_path
from vine_amd import autotune as AT
from vine_amd import learn as L
from vine_amd import solver as S
from vine_amd import workload as W


def _solve(g, src, dst, bound, policy=S.DELTA_MLP):
    opts = AT.bench_base(len(src))
    opts.update(delta_policy=policy, delta_fixed=4.0 if policy == S.DELTA_MLP else 12.0, concurrency=3)
    with S.Solver(g, **opts) as s:
        if policy == S.DELTA_MLP:
            s.set_mlp(np.load(L.DEFAULT_MODEL))
        return s.solve(src, dst, bound)


def _renumbered(g, perm):
    """The graph with vertex u renamed perm[u]; each vertex's arcs reversed."""
    n = g.n
    inv = np.argsort(perm)
    deg = np.diff(g.row_ptr)
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum(deg[inv])
    dst = np.empty(g.m, np.int32)
    ln = np.empty(g.m, np.float32)
    dl = np.empty(g.m, np.float32)
    for new_u in range(n):
        u = inv[new_u]
        lo, hi = g.row_ptr[u], g.row_ptr[u + 1]
        a, b = row[new_u], row[new_u + 1]
        dst[a:b] = perm[g.dst[lo:hi][::-1]]
        ln[a:b] = g.len[lo:hi][::-1]
        dl[a:b] = g.delay[lo:hi][::-1]
    return W.from_csr(row, dst, ln, dl)


@pytest.mark.gpu
def test_batch_metamorphic(gpu):
    g, q = W.preset("ny_batch")
    print("ny_batch: n=%d m=%d, %d queries, learned window" % (g.n, g.m, len(q)), flush=True)
    a = _solve(g, q.src, q.dst, q.bound)
    assert a.feasible.sum() > 0.9 * len(q)
    for i in range(len(q)):
        if a.feasible[i]:
            check_path(g, a.path_of(i), q.src[i], q.dst[i], a.cost[i], a.delay[i], q.bound[i])
        else:
            assert np.isinf(a.cost[i])

    print("lengths doubled", flush=True)
    g2 = W.from_csr(g.row_ptr, g.dst, g.len * np.float32(2), g.delay)
    b = _solve(g2, q.src, q.dst, q.bound)
    assert np.array_equal(b.feasible, a.feasible)
    f = a.feasible
    assert np.array_equal(b.cost[f], a.cost[f] * np.float32(2)) and np.array_equal(b.delay[f], a.delay[f])

    print("vertices renumbered, fixed window 12", flush=True)
    perm = np.random.default_rng(11).permutation(g.n).astype(np.int32)
    g3 = _renumbered(g, perm)
    c = _solve(g3, perm[q.src], perm[q.dst], q.bound, policy=S.DELTA_FIXED)
    assert np.array_equal(c.feasible, a.feasible)
    assert np.array_equal(c.cost[f], a.cost[f]) and np.array_equal(c.delay[f], a.delay[f])
    for i in range(0, len(q), 64):
        if c.feasible[i]:
            check_path(g3, c.path_of(i), perm[q.src[i]], perm[q.dst[i]], c.cost[i], c.delay[i], q.bound[i])

    print("bounds loosened by 10%", flush=True)
    looser = np.floor(q.bound * np.float32(1.1)).astype(np.float32)
    d = _solve(g, q.src, q.dst, looser)
    assert np.all(d.feasible[f])
    assert np.all(d.cost[f] <= a.cost[f])
