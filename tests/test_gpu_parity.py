"""Parity of the HIP solver with the CPU oracle (bit-exact costs, delays,
feasibility; valid paths; in the reference's one-to-all mode also the full
Pareto set at t)."""
import numpy as np
import pytest

from helpers import check_path
from oracle import oracle as O
from vine_amd import solver as S
from vine_amd import workload as W


@pytest.fixture(scope="module")
def grid_case():
    g, q = W.preset("grid64")
    ref = [O.cspp(g, q.src[i], q.dst[i], q.bound[i]) for i in range(len(q))]
    return g, q, ref


def _compare(g, q, ref, ans, idx=None):
    idx = range(len(q)) if idx is None else idx
    for k, i in enumerate(idx):
        r = ref[i]
        assert bool(ans.feasible[k]) == r.feasible, (i, ans.feasible[k], r.feasible)
        if r.feasible:
            assert ans.cost[k] == np.float32(r.cost), (i, ans.cost[k], r.cost)
            assert ans.delay[k] == np.float32(r.delay), (i, ans.delay[k], r.delay)
            check_path(g, ans.path_of(k), q.src[i], q.dst[i], ans.cost[k], ans.delay[k], q.bound[i])
        else:
            assert np.isinf(ans.cost[k])


@pytest.mark.gpu
@pytest.mark.parametrize("prune", [0, 1, 2])
@pytest.mark.parametrize("batch", [1, 100])
def test_grid_parity(gpu, grid_case, prune, batch):
    g, q, ref = grid_case
    print("grid64 parity prune=%d batch=%d" % (prune, batch), flush=True)
    with S.Solver(g, max_batch=batch, prune=prune) as s:
        ans = s.solve(q.src, q.dst, q.bound)
        _compare(g, q, ref, ans)
        if prune == 0 and batch == 100:
            # reference semantics: the full Pareto set at t is the oracle's
            for i in range(len(q)):
                got = s.pareto(i, q.dst[i])
                exp = ref[i].pareto
                assert got.shape == exp.shape, (i, got.shape, exp.shape)
                assert np.array_equal(got, exp), i


@pytest.mark.gpu
@pytest.mark.parametrize("policy,delta,engine", [(S.DELTA_FIXED, 0.0, S.ENGINE_ROUNDS),
                                                 (S.DELTA_FIXED, 2000.0, S.ENGINE_ROUNDS),
                                                 (S.DELTA_AR, 0.0, S.ENGINE_AUTO),
                                                 (S.DELTA_DOP, 0.0, S.ENGINE_AUTO),
                                                 (S.DELTA_FIXED, 0.0, S.ENGINE_WORKGROUP),
                                                 (S.DELTA_FIXED, 300.0, S.ENGINE_WORKGROUP),
                                                 (S.DELTA_FIXED, 2000.0, S.ENGINE_WORKGROUP)])
def test_grid_windows(gpu, grid_case, policy, delta, engine):
    g, q, ref = grid_case
    print("grid64 windows policy=%d delta=%g engine=%d" % (policy, delta, engine), flush=True)
    for batch in (1, 25):
        with S.Solver(g, max_batch=batch, prune=2, delta_policy=policy, delta_fixed=delta, engine=engine) as s:
            ans = s.solve(q.src, q.dst, q.bound)
        _compare(g, q, ref, ans)


def _tiny_graphs(count, seed):
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
        row = np.cumsum(row).astype(np.int32)
        ln = rng.integers(0, 10, u.shape[0]).astype(np.float32)
        dl = rng.integers(1, 10, u.shape[0]).astype(np.float32)
        yield W.Graph(row, v.astype(np.int32), ln, dl, "tiny")


@pytest.mark.gpu
def test_tiny_bruteforce(gpu):
    """Random tiny graphs (parallel arcs, zero lengths, cycles): every prune mode
    against exhaustive enumeration of simple paths."""
    rng = np.random.default_rng(1)
    for gi, g in enumerate(_tiny_graphs(40, 7)):
        nq = 12
        src = rng.integers(0, g.n, nq).astype(np.int32)
        dst = rng.integers(0, g.n, nq).astype(np.int32)
        bound = rng.integers(0, 25, nq).astype(np.float32)
        exp = [O.bruteforce(g, src[i], dst[i], bound[i]) for i in range(nq)]
        for prune in (0, 1, 2):
            with S.Solver(g, max_batch=5, prune=prune) as s:
                ans = s.solve(src, dst, bound)
            for i in range(nq):
                f, c, d = exp[i]
                assert bool(ans.feasible[i]) == f, (gi, prune, i)
                if f:
                    assert (float(ans.cost[i]), float(ans.delay[i])) == (c, d), (gi, prune, i)
                    check_path(g, ans.path_of(i), src[i], dst[i], ans.cost[i], ans.delay[i], bound[i])


@pytest.fixture(scope="module")
def rmat_case():
    g = W.rmat(scale=12, edge_factor=8)
    q = W.random_queries(g, 40, (13, 10), lcc_only=True, seed=3)
    ref = [O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2) for i in range(len(q))]
    return g, q, ref


@pytest.mark.gpu
@pytest.mark.parametrize("prune", [0, 2])
def test_rmat_hubs(gpu, rmat_case, prune):
    """Power-law graph: hub vertices with hundreds of arcs (the reference's
    `break` on an infeasible arc, algorithm/bucket.cuh:76-78, would skip
    strided arcs here)."""
    g, q, ref = rmat_case
    assert np.diff(g.row_ptr).max() > 64
    with S.Solver(g, max_batch=8, prune=prune) as s:
        ans = s.solve(q.src, q.dst, q.bound)
    _compare(g, q, ref, ans)


@pytest.mark.gpu
@pytest.mark.parametrize("bound_ball", [0, 1])
def test_knn_sample_bounds(gpu, bound_ball):
    """Geometric k-NN graph (the road-like generator) at 20k vertices with the
    lower-bound pruning, whole-frontier and windowed schedules; bounds from the
    forward-restricted searches (0) and from whole balls around t (1)."""
    g = W.knn(20000, seed=5)
    q = W.hop_queries(g, 30, hop_lo=5, hop_hi=80, f_permille=(1050, 1500), seed=5)
    ref = [O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2) for i in range(len(q))]
    print("knn20k bounds bound_ball=%d" % bound_ball, flush=True)
    for pol, dlt, eng in ((S.DELTA_INF, 0.0, S.ENGINE_AUTO), (S.DELTA_FIXED, 500.0, S.ENGINE_ROUNDS),
                          (S.DELTA_FIXED, 20.0, S.ENGINE_WORKGROUP), (S.DELTA_FIXED, 500.0, S.ENGINE_WORKGROUP)):
        with S.Solver(g, max_batch=16, prune=2, delta_policy=pol, delta_fixed=dlt, bound_ball=bound_ball,
                      engine=eng) as s:
            ans = s.solve(q.src, q.dst, q.bound)
        _compare(g, q, ref, ans)


@pytest.mark.gpu
def test_cpp_facade_matches_oracle(gpu):
    """The C++ facade's answers and paths (4x4 grid built in
    tests/cpp/facade_main.cpp) against the oracle on the same arcs."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([os.path.join(root, "vine_amd", "vine_facade_check")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [list(map(float, ln.split())) for ln in r.stdout.strip().splitlines()]
    from helpers import facade_grid
    g = facade_grid()
    tasks = [(0, 15, 100.0), (0, 15, 7.0), (5, 10, 2.0), (3, 3, 0.0)]
    for (s, t, b), out in zip(tasks, rows):
        ref = O.cspp(g, s, t, b)
        assert bool(out[0]) == ref.feasible
        if ref.feasible:
            assert (out[1], out[2]) == (ref.cost, ref.delay)
            check_path(g, np.array(out[3:], np.int64), s, t, out[1], out[2], b)


@pytest.mark.gpu
def test_workgroup_engine_eligibility(gpu, grid_case):
    """The workgroup engine needs the bounds and a finite window; asking for it
    otherwise is an error, and AUTO falls back to rounds."""
    g, q, ref = grid_case
    for kw in ({"prune": 1, "delta_policy": S.DELTA_FIXED, "delta_fixed": 10.0},
               {"prune": 2, "delta_policy": S.DELTA_INF},
               {"prune": 2, "delta_policy": S.DELTA_AR}):
        with pytest.raises(S.VineError):
            S.Solver(g, max_batch=4, engine=S.ENGINE_WORKGROUP, **kw)
    with S.Solver(g, max_batch=4, prune=1, delta_policy=S.DELTA_FIXED, delta_fixed=10.0) as s:
        ans = s.solve(q.src[:4], q.dst[:4], q.bound[:4])
    _compare(g, q, ref, ans, idx=range(4))


@pytest.mark.gpu
@pytest.mark.parametrize("block", [0, 32, 96, 320, 512, 1024])
def test_block_size_refused(gpu, grid_case, block):
    """The round kernels keep per-wave LDS tables for at most 256 threads (four
    waves) per workgroup: a larger workgroup, or one that is not whole waves,
    is refused with VINE_E_ARG when the solver is created, before any launch."""
    g, q, ref = grid_case
    print("block_size %d refused" % block, flush=True)
    with pytest.raises(S.VineError) as e:
        S.Solver(g, max_batch=4, block_size=block)
    assert e.value.status == S.VINE_E_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("block", [64, 192, 256])
def test_block_sizes_accepted(gpu, grid_case, block):
    g, q, ref = grid_case
    print("block_size %d" % block, flush=True)
    with S.Solver(g, max_batch=25, prune=2, delta_policy=S.DELTA_FIXED, delta_fixed=300.0, block_size=block,
                  engine=S.ENGINE_ROUNDS) as s:
        ans = s.solve(q.src, q.dst, q.bound)
    _compare(g, q, ref, ans)


def _fractional_knn(n, seed):
    """k-NN graph with weights that are multiples of 0.1 (not exact in binary):
    float sums then depend on their order."""
    g = W.knn(n, seed=seed)
    rng = np.random.default_rng(seed)
    ln = (rng.integers(1, 100, g.m) * 0.1).astype(np.float32)
    dl = (rng.integers(1, 50, g.m) * 0.1).astype(np.float32)
    return W.from_csr(g.row_ptr, g.dst, ln, dl, "knn-fractional")


@pytest.mark.gpu
def test_fractional_weights(gpu):
    """With weights that are not integers the bound pruning would add a label's
    delays in another order than the label (0.1 + (0.5 + 0.1) > (0.1 + 0.1) +
    0.5 in float32) and could drop the only feasible path; such graphs run
    target pruning, which compares a label's own sums only.  Answers against the
    oracle's reference search (one-to-all, no pruning) on the same floats."""
    # the path 0 -> 1 -> 2 -> 3 whose delay sum equals the bound exactly
    row = np.array([0, 1, 2, 3, 3], np.int32)
    dst = np.array([1, 2, 3], np.int32)
    ln = np.ones(3, np.float32)
    dl = np.array([0.1, 0.1, 0.5], np.float32)
    g = W.from_csr(row, dst, ln, dl, "path")
    bound = np.float32(np.float32(0.1) + np.float32(0.1)) + np.float32(0.5)
    assert np.float32(0.1) + np.float32(np.float32(0.5) + np.float32(0.1)) > bound
    ref = O.cspp(g, 0, 3, bound)
    assert ref.feasible and ref.delay == bound
    for prune in (S.PRUNE_TARGET, S.PRUNE_BOUNDS):
        with S.Solver(g, max_batch=1, prune=prune) as s:
            ans = s.solve(np.array([0], np.int32), np.array([3], np.int32), np.array([bound], np.float32))
        assert ans.feasible[0] and ans.cost[0] == 3 and ans.delay[0] == bound, prune
    # random queries on a geometric graph, every pruning mode and a window
    g = _fractional_knn(1500, 4)
    rng = np.random.default_rng(4)
    nq = 24
    src = rng.integers(0, g.n, nq).astype(np.int32)
    dst = rng.integers(0, g.n, nq).astype(np.int32)
    bound = np.zeros(nq, np.float32)
    for i in range(nq):
        md = O.dijkstra_lex(g, src[i], "delay", "len")[0][dst[i]]
        bound[i] = np.float32(1.2 * md) if np.isfinite(md) else np.float32(1.0)
    print("fractional weights: %d queries on %d vertices" % (nq, g.n), flush=True)
    ref = [O.cspp(g, src[i], dst[i], bound[i]) for i in range(nq)]
    q = W.Queries(src, dst, bound, None)
    for kw in ({"prune": S.PRUNE_BOUNDS}, {"prune": S.PRUNE_TARGET},
               {"prune": S.PRUNE_BOUNDS, "delta_policy": S.DELTA_FIXED, "delta_fixed": 5.0}):
        with S.Solver(g, max_batch=8, **kw) as s:
            ans = s.solve(src, dst, bound)
        _compare(g, q, ref, ans)


@pytest.mark.gpu
@pytest.mark.parametrize("policy", [S.DELTA_FIXED, S.DELTA_MLP, S.DELTA_INF])
def test_graph_replay(gpu, grid_case, policy):
    """use_graph = 1: every host check replays the captured chunk of rounds.
    Several solves in a row on one handle (the graph is captured again when
    the batch size changes), each against the oracle and against direct
    launches, with the first round of every chunk timed through event nodes."""
    from vine_amd import learn as L
    g, q, ref = grid_case
    delta = {S.DELTA_FIXED: 300.0, S.DELTA_MLP: 4.0, S.DELTA_INF: 0.0}[policy]
    kw = dict(max_batch=25, prune=2, delta_policy=policy, delta_fixed=delta, rounds_per_sync=4, time_kernels=1,
              engine=S.ENGINE_ROUNDS)
    print("graph replay policy=%d" % policy, flush=True)
    with S.Solver(g, **kw) as d:
        if policy == S.DELTA_MLP:
            d.set_mlp(np.load(L.DEFAULT_MODEL))
        direct = d.solve(q.src, q.dst, q.bound)
    _compare(g, q, ref, direct)
    with S.Solver(g, use_graph=1, **kw) as s:
        if policy == S.DELTA_MLP:
            s.set_mlp(np.load(L.DEFAULT_MODEL))
        for rep in range(3):
            a = s.solve(q.src, q.dst, q.bound)
            _compare(g, q, ref, a)
            assert np.array_equal(a.cost, direct.cost) and np.array_equal(a.delay, direct.delay), rep
            assert a.stats["timed_rounds"] > 0 and a.stats["ms_expand"] > 0, rep
        a = s.solve(q.src[:10], q.dst[:10], q.bound[:10])
        _compare(g, q, ref, a, idx=range(10))
        a = s.solve(q.src, q.dst, q.bound)
        _compare(g, q, ref, a)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2])
def test_graph_replay_streams(gpu, mode):
    """use_graph with three internal solvers (each captures on its own stream
    from its own host thread), on the geometric sample, against the oracle;
    mode 2 replays only once the frontier is small."""
    g = W.knn(20000, seed=5)
    q = W.hop_queries(g, 30, hop_lo=5, hop_hi=80, f_permille=(1050, 1500), seed=5)
    ref = [O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2) for i in range(len(q))]
    print("graph replay mode %d, 3 streams" % mode, flush=True)
    with S.Solver(g, max_batch=30, prune=2, delta_policy=S.DELTA_FIXED, delta_fixed=500.0, rounds_per_sync=8,
                  use_graph=mode, concurrency=3, engine=S.ENGINE_ROUNDS) as s:
        for rep in range(2):
            ans = s.solve(q.src, q.dst, q.bound)
            _compare(g, q, ref, ans)


@pytest.mark.gpu
@pytest.mark.parametrize("slots", [16, 128, 1024])
def test_lds_candidate_table(gpu, grid_case, slots):
    """The first-level candidate table in LDS (hash_slots): a tiny table sends
    most candidates to the overflow path (appended to HBM directly), a large
    one keeps them; answers equal the oracle's either way, with a fixed window
    and with the learned window that sizes the table each round."""
    from vine_amd import learn as L
    g, q, ref = grid_case
    print("LDS candidate table, %d slots" % slots, flush=True)
    for policy, delta in ((S.DELTA_FIXED, 300.0), (S.DELTA_MLP, 4.0), (S.DELTA_INF, 0.0)):
        with S.Solver(g, max_batch=25, prune=2, delta_policy=policy, delta_fixed=delta, hash_slots=slots,
                      engine=S.ENGINE_ROUNDS) as s:
            if policy == S.DELTA_MLP:
                s.set_mlp(np.load(L.DEFAULT_MODEL))
            ans = s.solve(q.src, q.dst, q.bound)
        _compare(g, q, ref, ans)
        st = ans.stats
        print("  policy %d: candidates %d, dropped %d, overflow %d" % (policy, st["candidates"], st["hash_dropped"],
                                                                      st["hash_overflow"]), flush=True)
        if slots == 16 and policy == S.DELTA_INF:
            assert st["hash_overflow"] > 0


@pytest.mark.gpu
def test_lds_candidate_table_knn(gpu):
    """The table on the geometric sample with the reference's one-to-all search
    (full Pareto set at t compared) and with the bounds."""
    g = W.knn(20000, seed=5)
    q = W.hop_queries(g, 30, hop_lo=5, hop_hi=80, f_permille=(1050, 1500), seed=5)
    ref = [O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2) for i in range(len(q))]
    print("LDS candidate table, knn20k", flush=True)
    for prune, slots in ((2, 64), (2, 1024), (1, 256)):
        with S.Solver(g, max_batch=16, prune=prune, delta_policy=S.DELTA_FIXED, delta_fixed=500.0,
                      hash_slots=slots, engine=S.ENGINE_ROUNDS) as s:
            ans = s.solve(q.src, q.dst, q.bound)
        _compare(g, q, ref, ans)
    with S.Solver(g, max_batch=30, prune=0, hash_slots=32) as s:
        ans = s.solve(q.src[:6], q.dst[:6], q.bound[:6])
        for i in range(6):
            full = O.cspp(g, q.src[i], q.dst[i], q.bound[i])
            got = s.pareto(i, q.dst[i])
            assert np.array_equal(got, full.pareto), i
    assert ans.stats["hash_dropped"] + ans.stats["hash_overflow"] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("max_labels", [2e4, 2e5, 0.0])
def test_dop_count_semantics(gpu, grid_case, max_labels):
    """The reference's DOP rule with its count semantics: a round expands the
    last DOP frontier entries and carries the rest.  A small MAX makes the
    count bind (more rounds than whole-frontier rounds); answers equal the
    oracle's either way."""
    g, q, ref = grid_case
    print("DOP count semantics, MAX %g" % max_labels, flush=True)
    with S.Solver(g, max_batch=25, prune=2, delta_policy=S.DELTA_INF) as s:
        whole = s.solve(q.src, q.dst, q.bound)
    with S.Solver(g, max_batch=25, prune=2, delta_policy=S.DELTA_DOP, ar_max_labels=max_labels) as s:
        ans = s.solve(q.src, q.dst, q.bound)
    _compare(g, q, ref, ans)
    if max_labels == 2e4:
        assert ans.stats["rounds"] > whole.stats["rounds"]


@pytest.mark.gpu
@pytest.mark.parametrize("lagrange", [0, 1, 2, 4])
@pytest.mark.parametrize("policy,delta", [(S.DELTA_FIXED, 200.0), (S.DELTA_INF, 0.0)])
def test_lagrange_bound_grid(gpu, grid_case, lagrange, policy, delta):
    """The bound pruning with and without the Lagrangian search: the same
    answers as the oracle on the grid workload (len 1-1000 against delay 1-100,
    so the multipliers are large)."""
    g, q, ref = grid_case
    print("grid64 lagrange=%d policy=%d" % (lagrange, policy), flush=True)
    with S.Solver(g, max_batch=100, prune=2, delta_policy=policy, delta_fixed=delta, lagrange_bound=lagrange) as s:
        ans = s.solve(q.src, q.dst, q.bound)
    _compare(g, q, ref, ans)


@pytest.mark.gpu
def test_lagrange_bound_tight_knn(gpu):
    """Tight bounds (1.02-1.10 x the minimum delay) on a 20k-vertex road-like
    graph, where the Lagrangian bound prunes most: answers and paths equal the
    label-setting oracle's with and without it, and it expands fewer labels."""
    g = W.knn(20000, seed=11)
    q = W.hop_queries(g, 48, hop_lo=20, hop_hi=120, f_permille=(1020, 1100), seed=11)
    ref = [O.cspp_astar(g, q.src[i], q.dst[i], q.bound[i]) for i in range(len(q))]
    expanded = {}
    for lagrange in (0, 2, 4):
        print("knn20k tight lagrange=%d" % lagrange, flush=True)
        with S.Solver(g, max_batch=48, prune=2, delta_policy=S.DELTA_FIXED, delta_fixed=40.0,
                      lagrange_bound=lagrange) as s:
            ans = s.solve(q.src, q.dst, q.bound)
        _compare(g, q, ref, ans)
        expanded[lagrange] = ans.stats["expanded"]
    assert expanded[2] < expanded[0] and expanded[4] < expanded[0], expanded


@pytest.mark.gpu
def test_lagrange_bound_large_weights(gpu):
    """Lengths up to 100,000 per arc: the Lagrangian sums (256 x length plus k x
    delay, unsigned integers) reach hundreds of millions, far past where float
    sums stay exact, while the plain sums stay below 2^24; every answer is
    still the oracle's."""
    g0, q = W.preset("grid64")
    g = W.Graph(g0.row_ptr, g0.dst, (g0.len * 100.0).astype(np.float32), g0.delay)
    ref = [O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2) for i in range(len(q))]
    assert max(r.cost for r in ref if r.feasible) * 256 > 2 ** 28
    with S.Solver(g, max_batch=100, prune=2, delta_policy=S.DELTA_FIXED, delta_fixed=20000.0, lagrange_bound=2) as s:
        ans = s.solve(q.src, q.dst, q.bound)
    _compare(g, q, ref, ans)


@pytest.mark.gpu
def test_lagrange_bound_refused_values(gpu, grid_case):
    g, q, _ = grid_case
    with pytest.raises(S.VineError):
        S.Solver(g, max_batch=4, prune=2, lagrange_bound=9)
