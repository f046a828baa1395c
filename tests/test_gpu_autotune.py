"""The launch-shape autotuner on the GPU: a full coordinate search over its
space on a small geometric workload (every configuration must give the same
answers, which are also the oracle's), and the cache round trip."""
import json

import numpy as np
import pytest

from oracle import oracle as O
from vine_amd import autotune as AT
from vine_amd import learn as L
from vine_amd import solver as S
from vine_amd import workload as W


@pytest.mark.gpu
def test_autotune_search_and_cache(gpu, tmp_path):
    g = W.knn(20000, seed=6)
    q = W.hop_queries(g, 96, hop_lo=5, hop_hi=120, f_permille=(1050, 2000), seed=6)
    dg = S.DeviceGraph(g)
    cache = str(tmp_path / "autotune.json")
    base = dict(max_batch=96, prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_MLP, delta_fixed=4.0, rounds_per_sync=8)
    print("autotune search, %d queries" % len(q), flush=True)
    opts = AT.tuned_options(dg, g, q, cache=cache, mlp=np.load(L.DEFAULT_MODEL), log=print, **base)
    for names, vals in AT.SPACE:
        assert tuple(opts[k] for k in names) in vals, names
    entry = json.load(open(cache))[AT.graph_key(g, 96, S.DELTA_MLP)]
    assert len(entry["tried"]) == sum(len(v) for names, v in AT.SPACE if AT.applies(names, opts))
    assert AT.cached_options(g, cache=cache, **base) == {k: opts[k] for k in AT.KEYS}
    # the chosen configuration's answers are the oracle's
    with S.Solver(dg, **opts) as s:
        s.set_mlp(np.load(L.DEFAULT_MODEL))
        a = s.solve(q.src, q.dst, q.bound)
    for i in range(0, len(q), 4):
        r = O.cspp_astar(g, q.src[i], q.dst[i], q.bound[i])
        assert bool(a.feasible[i]) == r.feasible, i
        if r.feasible:
            assert (a.cost[i], a.delay[i]) == (np.float32(r.cost), np.float32(r.delay)), i
