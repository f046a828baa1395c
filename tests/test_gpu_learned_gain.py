"""What the trained window network gains over the method without it.

The method without it is the reference's bucket-width rule: the DOP count of
include/AR.hpp driven by the AR(3) forecast, with the reference's MAX (free
device memory / sizeof(label), algorithm/cspp.cuh:43-47).  Both runs use the
same kernels, bounds and batch; only the window policy differs, so answers
must be identical and the gain shows in labels expanded, rounds and time.
Both run the grid-wide round engine, the only one the AR rule runs in.
The network is the committed one (vine_amd/models/delta_mlp.npy), trained on
NY-scale queries with seeds disjoint from these.
"""
import time

import numpy as np
import pytest

from vine_amd import learn as L
from vine_amd import solver as S
from vine_amd import workload as W


def _run(dg, q, policy, delta, params=None):
    # every kernel timed with HIP events: the AR rule reads the frontier size
    # on the host after every round, the learned window every 16 rounds, so
    # wall time would also count those waits; kernel time does not
    with S.Solver(dg, max_batch=len(q), prune=S.PRUNE_BOUNDS, delta_policy=policy, delta_fixed=delta,
                  rounds_per_sync=16, time_kernels=1, engine=S.ENGINE_ROUNDS) as s:
        if params is not None:
            s.set_mlp(params)
        s.solve(q.src[:8], q.dst[:8], q.bound[:8], paths=False)  # warm-up (code objects, allocations)
        t0 = time.perf_counter()
        a = s.solve(q.src, q.dst, q.bound, paths=False)
        return a, time.perf_counter() - t0


@pytest.mark.gpu
def test_learned_window_beats_ar_rule(gpu):
    params = np.load(L.DEFAULT_MODEL)
    g = W.knn(50000, seed=7)
    q = W.hop_queries(g, 256, hop_lo=20, hop_hi=300, seed=7)
    dg = S.DeviceGraph(g)
    ar, t_ar = _run(dg, q, S.DELTA_AR, 0.0)
    ml, t_ml = _run(dg, q, S.DELTA_MLP, 4.0, params)
    cost = {}
    kernel_ms = {}
    for name, a, t in (("ar", ar, t_ar), ("mlp", ml, t_ml)):
        st = a.stats
        cost[name] = st["expanded"] + L.ROUND_COST * st["rounds"]
        kernel_ms[name] = st["ms_expand"] + st["ms_merge"] + st["ms_pull"] + st["ms_sssp"]
        print("%s: %.3f s wall, %.1f ms in kernels, %d rounds, %d labels expanded, modelled cost %.3g"
              % (name, t, kernel_ms[name], st["rounds"], st["expanded"], cost[name]), flush=True)
    # scheduling only: the same exact answers
    assert np.array_equal(ar.feasible, ml.feasible)
    assert np.array_equal(ar.cost, ml.cost) and np.array_equal(ar.delay, ml.delay)
    assert ml.feasible.sum() > 0.9 * len(q)
    # the gain: less work in the measured cost model, and less GPU time
    assert ml.stats["expanded"] < 0.5 * ar.stats["expanded"]
    assert cost["mlp"] < 0.5 * cost["ar"]
    assert kernel_ms["mlp"] < kernel_ms["ar"]
