"""The generic label-family search (vine_family_solve) on the GPU against the
family oracle: bit-exact best labels and whole Pareto sets at t, and every
returned path re-walked arc by arc to the returned label with the constraint
holding on each prefix.  Batches smaller than the query set exercise the
batch loop; c_label through the generic search must also match the main
solver."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

from oracle import oracle as O
from test_oracle_families import passes, path_label
from vine_amd import solver as S
from vine_amd import workload as W

F32 = np.float32


def _social_graph(seed):
    g = W.grid(20, 20, seed=seed)
    rng = np.random.default_rng(seed)
    g.len = (rng.integers(1, 17, g.m) / 16.0).astype(F32)
    g.delay = (rng.integers(1, 17, g.m) / 16.0).astype(F32)
    return g


def _queries(g, fam, nq, seed):
    rng = np.random.default_rng(seed)
    q = W.random_queries(g, nq, (13, 10), seed=seed)
    src_idx = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    hop_d = dijkstra(sp.csr_matrix((np.ones(g.m), (src_idx, g.dst)), shape=(g.n, g.n)), indices=q.src)
    hops = np.array([hop_d[i, q.dst[i]] for i in range(nq)])
    hops[np.isinf(hops)] = 0
    lim = np.zeros((nq, 3), np.float32)
    lim[:, 0] = q.bound
    lim[:, 1] = hops + rng.integers(0, 8, nq)
    lim[:, 2] = rng.integers(400, 1001, nq)
    if fam == S.FAMILY_QOS:
        lim[:, 0] = rng.integers(1, 60, nq)
    if fam == S.FAMILY_SOCIAL:
        lim[:, 0] = 1.0
    return q.src, q.dst, lim


def _arc(g, a, b):
    lo, hi = g.row_ptr[a], g.row_ptr[a + 1]
    ks = np.nonzero(g.dst[lo:hi] == b)[0]
    assert ks.size == 1, (a, b)
    return lo + ks[0]


def _check_path(g, fam, path, s, t, label, lim):
    assert path[0] == s and path[-1] == t
    arcs = [_arc(g, a, b) for a, b in zip(path[:-1], path[1:])]
    labs = path_label(fam, g, arcs)
    assert all(passes(fam, lab, lim) for lab in labs)
    if labs:
        assert np.array_equal(np.array(labs[-1], F32), label)


@pytest.mark.gpu
@pytest.mark.parametrize("prune,window", [(0, 0.0), (1, 0.0), (1, 400.0)])
@pytest.mark.parametrize("fam", [S.FAMILY_CSP, S.FAMILY_HOP, S.FAMILY_HOP_MAXCOST, S.FAMILY_2OSP, S.FAMILY_SOCIAL,
                                 S.FAMILY_QOS])
def test_family_parity(gpu, fam, prune, window):
    g = _social_graph(3) if fam == S.FAMILY_SOCIAL else W.grid(24, 24, seed=fam)
    src, dst, lim = _queries(g, fam, 48, seed=fam)
    print("family %d, prune %d, window %g: n=%d, %d queries in batches of 16" % (fam, prune, window, g.n, len(src)),
          flush=True)
    with S.FamilySolver(g, fam, max_batch=16, prune=prune, window=window) as fs:
        a = fs.solve(src, dst, lim[:, 0], lim[:, 1], lim[:, 2])
        refs = [O.family(g, fam, src[i], dst[i], lim[i]) for i in range(len(src))]
        for i, r in enumerate(refs):
            assert bool(a.feasible[i]) == r.feasible, (i, a.label[i], r.best)
            if r.feasible:
                assert np.array_equal(a.label[i], r.best), (i, a.label[i], r.best)
                _check_path(g, fam, a.path_of(i), src[i], dst[i], a.label[i], lim[i])
            else:
                assert np.all(np.isinf(a.label[i]))
        # whole Pareto sets at t for the last batch (queries 32..47); target
        # pruning keeps only the answer exact
        for k in range(16 if prune == 0 else 0):
            i = 32 + k
            got = fs.front(k, dst[i])
            order = np.lexsort(tuple(got[:, j] for j in range(3, -1, -1)))
            exp = refs[i].front
            eorder = np.lexsort(tuple(exp[:, j] for j in range(3, -1, -1)))
            assert np.array_equal(got[order], exp[eorder]), i
    assert a.counters["rounds"] > 0 and a.counters["inserted"] > 0


@pytest.mark.gpu
def test_social_tight_bound_paths_are_valid(gpu):
    """delay_L < 1: the answer depends on the order labels meet (the family's
    constraint runs against its dominance), so only validity is checked."""
    g = _social_graph(5)
    rng = np.random.default_rng(5)
    src, dst = rng.integers(0, g.n, 40).astype(np.int32), rng.integers(0, g.n, 40).astype(np.int32)
    lim = np.zeros((40, 3), np.float32)
    lim[:, 0] = rng.choice([0.5, 0.25, 0.125], 40)
    print("social, tight bounds: %d queries" % len(src), flush=True)
    with S.FamilySolver(g, S.FAMILY_SOCIAL, max_batch=40) as fs:
        a = fs.solve(src, dst, lim[:, 0])
    for i in range(40):
        if a.feasible[i] and src[i] != dst[i]:
            _check_path(g, S.FAMILY_SOCIAL, a.path_of(i), src[i], dst[i], a.label[i], lim[i])


@pytest.mark.gpu
def test_c_label_family_matches_main_solver(gpu):
    g, q = W.preset("grid64")
    print("c_label through the generic search vs the main solver: %d queries" % len(q), flush=True)
    with S.FamilySolver(g, S.FAMILY_CSP, max_batch=100) as fs:
        a = fs.solve(q.src, q.dst, q.bound)
    with S.Solver(g, max_batch=100, prune=S.PRUNE_BOUNDS) as s:
        b = s.solve(q.src, q.dst, q.bound)
    assert np.array_equal(a.feasible, b.feasible)
    assert np.array_equal(a.label[:, 0], b.cost) and np.array_equal(a.label[:, 1], b.delay)


@pytest.mark.gpu
def test_family_argument_errors(gpu):
    g = W.grid(8, 8, seed=1)
    with pytest.raises(S.VineError, match="social"):
        S.FamilySolver(g, S.FAMILY_SOCIAL)
    with pytest.raises(S.VineError):
        S.FamilySolver(g, 9)
    with S.FamilySolver(g, S.FAMILY_HOP, max_batch=4, label_capacity=16) as fs:
        with pytest.raises(S.VineError, match="vertex id"):
            fs.solve(np.array([0], np.int32), np.array([64], np.int32), np.array([10.0], np.float32))
        with pytest.raises(S.VineError, match="capacity"):
            fs.solve(np.array([0], np.int32), np.array([63], np.int32), np.array([1e6], np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("mode,fam,limits", [
    ("hop", S.FAMILY_HOP, [(100, 6, 0), (100, 5, 0), (4, 100, 0), (0, 100, 0)]),
    ("qos", S.FAMILY_QOS, [(1, 0, 0), (2, 0, 0), (3, 0, 0), (1, 0, 0)]),
])
def test_cpp_facade_families(gpu, mode, fam, limits):
    """cspp<hop::c_label> and cspp<qos::c_label> of the C++ facade (the reference's
    solver API with another label family) on the grid of tests/cpp/facade_main.cpp."""
    import os
    import subprocess

    from helpers import facade_grid
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("facade, family %s" % mode, flush=True)
    r = subprocess.run([os.path.join(root, "vine_amd", "vine_facade_check"), mode], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [ln.split() for ln in r.stdout.strip().splitlines()]
    g = facade_grid()
    tasks = [(0, 15), (0, 15), (5, 10), (3, 3)]
    for (s, t), lim, out in zip(tasks, limits, rows):
        ref = O.family(g, fam, s, t, lim)
        assert bool(int(out[0])) == ref.feasible, (mode, s, t, out)
        if ref.feasible:
            assert [float(x) for x in out[1:4]] == [float(x) for x in ref.best[:3]], (mode, s, t, out, ref.best)
            label = np.array([float(x) for x in out[1:4]] + [0.0], F32)
            _check_path(g, fam, np.array([int(x) for x in out[4:]]), s, t, label, np.array(lim, F32))
