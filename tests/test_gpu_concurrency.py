"""vine_options.concurrency: one handle splitting every solve over several
internal solvers on their own HIP streams.  Answers, paths, per-query records
and Pareto sets must equal those of a single-stream handle, through both
entry points (host buffers with paths, device buffers)."""
import numpy as np
import pytest
import torch

from helpers import check_path
from vine_amd import solver as S
from vine_amd import workload as W


def _same(a, b):
    assert np.array_equal(a.cost, b.cost) and np.array_equal(a.delay, b.delay)
    assert np.array_equal(a.feasible, b.feasible)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [2, 3, 5])
def test_split_matches_single_stream(gpu, k):
    g, q = W.preset("grid64")
    print("grid64, concurrency %d" % k, flush=True)
    opts = dict(max_batch=40, prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_FIXED, delta_fixed=300.0,
                rounds_per_sync=8)
    with S.Solver(g, **opts) as s1, S.Solver(g, concurrency=k, **opts) as sk:
        a = s1.solve(q.src, q.dst, q.bound)
        b = sk.solve(q.src, q.dst, q.bound)
        _same(a, b)
        for i in range(len(q)):
            if b.feasible[i]:
                check_path(g, b.path_of(i), q.src[i], q.dst[i], b.cost[i], b.delay[i], q.bound[i])
        ex1, r1 = s1.query_stats(len(q))
        exk, rk = sk.query_stats(len(q))
        assert (exk > 0).sum() == (ex1 > 0).sum()
        assert np.all(sk.latency_us(len(q))[b.feasible] > 0)
        assert b.stats["batches"] >= k and b.stats["expanded"] > 0


@pytest.mark.gpu
def test_split_device_entry_and_pareto(gpu):
    g, q = W.preset("grid64")
    print("grid64, concurrency 3, device buffers", flush=True)
    opts = dict(max_batch=100, prune=S.PRUNE_NONE, delta_policy=S.DELTA_FIXED, delta_fixed=500.0)
    dev = torch.device("cuda:0")
    with S.Solver(g, **opts) as s1, S.Solver(g, concurrency=3, **opts) as s3:
        a = s1.solve(q.src, q.dst, q.bound, paths=False)
        ts = [torch.from_numpy(x).to(dev) for x in (q.src, q.dst, q.bound)]
        cost = torch.empty(len(q), dtype=torch.float32, device=dev)
        delay = torch.empty(len(q), dtype=torch.float32, device=dev)
        feas = torch.empty(len(q), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        st = s3.solve_device(ts[0].data_ptr(), ts[1].data_ptr(), ts[2].data_ptr(), cost.data_ptr(),
                             delay.data_ptr(), feas.data_ptr(), len(q))
        assert np.array_equal(cost.cpu().numpy(), a.cost) and np.array_equal(delay.cpu().numpy(), a.delay)
        assert np.array_equal(feas.cpu().numpy().astype(bool), a.feasible)
        assert st["rounds"] > 0
        # the reference's one-to-all Pareto sets, query by query
        for i in range(0, len(q), 9):
            assert np.array_equal(s1.pareto(i, q.dst[i]), s3.pareto(i, q.dst[i])), i


@pytest.mark.gpu
def test_split_learned_window_on_ny_sample(gpu):
    g = W.knn(264346, seed=0)
    q = W.hop_queries(g, 300, seed=1)
    params = np.load(S.__file__.replace("solver.py", "models/delta_mlp.npy"))
    print("ny sample: 300 hop queries, learned window, concurrency 1 vs 3", flush=True)
    res = []
    for k in (1, 3):
        with S.Solver(g, max_batch=300, prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_MLP, delta_fixed=4.0,
                      rounds_per_sync=16, concurrency=k, label_capacity=60_000_000) as s:
            s.set_mlp(params)
            res.append(s.solve(q.src, q.dst, q.bound, paths=False))
    _same(res[0], res[1])


@pytest.mark.gpu
def test_concurrency_option_range(gpu):
    g = W.grid(8, 8, seed=1)
    with pytest.raises(S.VineError, match="concurrency"):
        S.Solver(g, concurrency=17)
