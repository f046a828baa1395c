"""Full-size workloads on the GPU, every answer against the oracle's.

The answers come from tests/golden/answers_<workload>.npz, written offline by
tests/golden/make_full_answers.py (label-setting oracle for every query,
cross-checked by the restated cspp_cpu search on the queries it finishes).
Each test solves the fixture's own queries on the graph regenerated from the
same seed (its CRC must match) and compares cost, delay and feasibility bit
for bit; every returned path is re-walked (valid arcs, float32 sums equal to
the cost and delay, delay within the bound).

C5: NY-scale 3-NN graph (264,346 vertices), the 4,096-query batched workload,
    in the benchmark's configuration and with the plain single-query solver on
    a sample.
C2: the same graph, 1,000 random pairs, bounds 1.1x and 1.5x.
C3: NE-scale 3-NN graph (1,524,453 vertices), 1,000 random pairs, bound 1.2x,
    learned window against a fixed one.
C4: R-MAT scale 17 (131,072 vertices, hubs of hundreds of arcs), 500 pairs
    from the largest component, bound 1.3x.
"""
import numpy as np
import pytest

from helpers import check_paths_all, compare_fixture, fixture_answers, graph_crc
from vine_amd import autotune as AT
from vine_amd import learn as L
from vine_amd import solver as S
from vine_amd import workload as W

NY_N, NE_N = 264346, 1524453


def _case(name):
    fx = fixture_answers(name)
    if name.startswith("ny"):
        g = W.knn(NY_N, seed=0)
    elif name.startswith("ne"):
        g = W.knn(NE_N, seed=0)
    else:
        g = W.rmat(seed=0)
    assert (g.n, g.m) == (int(fx["n"]), int(fx["m"])) and graph_crc(g) == int(fx["graph_crc"]), name
    q = W.Queries(fx["src"], fx["dst"], fx["bound"], None)
    return g, q, fx


def _solve(g, q, batch, mlp=None, **kw):
    opts = dict(prune=S.PRUNE_BOUNDS, rounds_per_sync=16)
    opts.update(kw)
    opts["max_batch"] = batch
    with S.Solver(g, **opts) as s:
        if mlp is not None:
            s.set_mlp(mlp)
        return s.solve(q.src, q.dst, q.bound)


def _check(name, g, q, fx, a, idx=None):
    n = compare_fixture(fx, a, idx)
    sub = q if idx is None else q.subset(idx)
    check_paths_all(g, sub.src, sub.dst, sub.bound, a)
    print("%s: %d answers bit-exact vs the oracle, paths valid" % (name, n), flush=True)


ENGINES = [S.ENGINE_ROUNDS, S.ENGINE_WORKGROUP]


@pytest.mark.gpu
def test_ny_batch_full(gpu):
    """The benchmark's workload and configuration: 4,096 queries, learned
    window, the cached launch options; then the learned window in the other
    engine and a fixed window in both, in one stream."""
    g, q, fx = _case("ny_batch")
    print("ny_batch: %d queries" % len(q), flush=True)
    base = AT.bench_base(4096)
    launch = AT.cached_options(g, **base) or {"concurrency": 3}
    dg = S.DeviceGraph(g)
    mlp = np.load(L.DEFAULT_MODEL)
    a = _solve(dg, q, 4096, mlp=mlp, **{**base, **launch})
    _check("ny_batch learned (bench options %s)" % launch, g, q, fx, a)
    other = S.ENGINE_ROUNDS if launch.get("engine", S.ENGINE_AUTO) != S.ENGINE_ROUNDS else S.ENGINE_WORKGROUP
    a2 = _solve(dg, q, 4096, mlp=mlp, **{**base, "engine": other})
    _check("ny_batch learned engine %d" % other, g, q, fx, a2)
    for eng in ENGINES:
        b = _solve(dg, q, 4096, delta_policy=S.DELTA_FIXED, delta_fixed=12.0, engine=eng, **AT.BENCH_CAPACITIES)
        _check("ny_batch fixed 12 engine %d" % eng, g, q, fx, b)


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ENGINES)
def test_ny_batch_plain_sample(gpu, engine):
    """The plain single-query solver (one query per solve) on 96 queries of the
    batched workload drawn across its hop range."""
    g, q, fx = _case("ny_batch")
    idx = np.random.default_rng(3).choice(len(q), 96, replace=False)
    sub = q.subset(idx)
    a = _solve(g, sub, 1, delta_policy=S.DELTA_FIXED, delta_fixed=12.0, label_capacity=200_000_000, engine=engine)
    _check("ny_batch plain engine %d" % engine, g, q, fx, a, idx=idx)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ny_1.1x", "ny_1.5x"])
@pytest.mark.parametrize("engine", ENGINES)
def test_ny_random_pairs_full(gpu, name, engine):
    g, q, fx = _case(name)
    print("%s: %d queries, engine %d" % (name, len(q), engine), flush=True)
    dg = S.DeviceGraph(g)
    a = _solve(dg, q, 1000, delta_policy=S.DELTA_FIXED, delta_fixed=12.0, engine=engine)
    _check(name + " fixed 12", g, q, fx, a)
    b = _solve(dg, q, 1000, mlp=np.load(L.DEFAULT_MODEL), delta_policy=S.DELTA_MLP, delta_fixed=4.0, concurrency=2,
               engine=engine)
    _check(name + " learned", g, q, fx, b)


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ENGINES)
def test_ne_full_learned_vs_fixed(gpu, engine):
    """All 1,000 NE pairs at 1.2x: the learned window against a fixed one, in
    batches of 250 query slots; the learned window expands fewer labels."""
    g, q, fx = _case("ne_1.2x")
    print("ne: n=%d m=%d, %d queries at 1.2x, engine %d" % (g.n, g.m, len(q), engine), flush=True)
    dg = S.DeviceGraph(g)
    a = _solve(dg, q, 250, delta_policy=S.DELTA_FIXED, delta_fixed=12.0, label_capacity=300_000_000, engine=engine)
    _check("ne fixed 12", g, q, fx, a)
    b = _solve(dg, q, 250, mlp=np.load(L.DEFAULT_MODEL), delta_policy=S.DELTA_MLP, delta_fixed=4.0,
               concurrency=3, label_capacity=300_000_000, engine=engine)
    _check("ne learned", g, q, fx, b)
    assert b.stats["expanded"] < a.stats["expanded"]


@pytest.mark.gpu
def test_rmat17_full(gpu):
    g, q, fx = _case("rmat17")
    print("rmat17: n=%d m=%d max out-degree %d, %d queries" % (g.n, g.m, np.diff(g.row_ptr).max(), len(q)),
          flush=True)
    assert fx["feasible"].all()  # pairs come from the largest component, bound >= min delay
    dg = S.DeviceGraph(g)
    # out-degrees above 512: the workgroup engine does not apply, the default picks grid-wide rounds
    with pytest.raises(S.VineError):
        S.Solver(dg, max_batch=4, prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_FIXED, delta_fixed=20.0,
                 engine=S.ENGINE_WORKGROUP)
    a = _solve(dg, q, 500, delta_policy=S.DELTA_FIXED, delta_fixed=20.0)
    _check("rmat17 fixed 20", g, q, fx, a)
    b = _solve(dg, q, 500, delta_policy=S.DELTA_INF)
    _check("rmat17 whole frontier", g, q, fx, b)
