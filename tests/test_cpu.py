"""CPU-side checks of the product library: the C ABI loads and exports every
symbol the headers declare, the workload generators are deterministic and have
the documented shape, and the host scheduling rule (AR / DOP) matches the
reference's values (golden vectors from the compiled include/AR.cpp)."""
import ctypes
import hashlib
import json
import os
import re

import numpy as np
import pytest

from vine_amd import _lib
from vine_amd import solver as S
from vine_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "ar_golden.json")


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(vine_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def _struct_fields(name):
    text = open(os.path.join(ROOT, "include", "vine.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct \{([^}]*)\}\s*%s;" % name, text).group(1)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(\w+)\s+(\w+)(\[(\d+)\])?$", decl)
        assert m, decl
        out.append((m.group(2), m.group(1), int(m.group(4) or 1)))
    return out


def test_options_layout_matches_header():
    o = S.default_options()
    assert o.max_batch == 1 and o.prune == S.PRUNE_BOUNDS and o.block_size == 128
    for cname, pystruct in (("vine_options", S.Options), ("vine_stats", S.Stats)):
        hdr = _struct_fields(cname)
        py = [(f[0], f[1]) for f in pystruct._fields_]
        assert [h[0] for h in hdr] == [p[0] for p in py], cname
        size = {"int32_t": 4, "int64_t": 8, "float": 4, "double": 8}
        for (nm, ty, cnt), (_, cty) in zip(hdr, py):
            assert ctypes.sizeof(cty) == size[ty] * cnt, (cname, nm)


def test_errors_without_device():
    # argument errors are reported before any device work
    L = S._L()
    h = ctypes.c_void_p()
    rc = L.vine_graph_create(0, 0, None, None, None, None, ctypes.byref(h))
    assert rc == S.VINE_E_ARG
    assert b"invalid" in L.vine_last_error()


def _h(*arrs):
    m = hashlib.sha256()
    for a in arrs:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()[:16]


def test_generators_deterministic():
    a = W.grid(seed=0)
    b = W.grid(seed=0)
    c = W.grid(seed=1)
    assert _h(a.row_ptr, a.dst, a.len, a.delay) == _h(b.row_ptr, b.dst, b.len, b.delay)
    assert _h(a.row_ptr, a.dst, a.len, a.delay) != _h(c.row_ptr, c.dst, c.len, c.delay)
    k1 = W.knn(5000, seed=3)
    k2 = W.knn(5000, seed=3)
    assert _h(k1.dst, k1.len) == _h(k2.dst, k2.len)
    q1 = W.hop_queries(k1, 50, hop_lo=3, hop_hi=40, seed=2, nthreads=1)
    q2 = W.hop_queries(k1, 50, hop_lo=3, hop_hi=40, seed=2, nthreads=4)
    assert np.array_equal(q1.src, q2.src) and np.array_equal(q1.bound, q2.bound)


def _symmetric(g):
    u = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    fwd = set(zip(u.tolist(), g.dst.tolist(), g.len.tolist(), g.delay.tolist()))
    return all((v, x, l, d) in fwd for (x, v, l, d) in fwd)


def test_grid_shape():
    g = W.grid()
    assert g.n == 4096 and g.m == 14516
    assert g.len.min() >= 1 and g.len.max() <= 1000 and g.delay.min() >= 1 and g.delay.max() <= 100
    assert np.all(g.len == np.round(g.len)) and np.all(g.delay == np.round(g.delay))
    assert _symmetric(g)


def test_knn_shape():
    g = W.knn(20000, seed=0)
    deg = np.diff(g.row_ptr)
    assert deg.min() >= 3  # every point keeps its 3 nearest neighbours
    assert g.len.min() >= 1 and np.all(g.delay >= 1)
    ratio = g.delay / g.len
    assert ratio.min() >= 0.5 - 0.5 / g.len.min() and ratio.max() <= 2.0 + 0.5
    assert _symmetric(g)


def test_rmat_shape():
    g = W.rmat(scale=10, edge_factor=8)
    u = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    assert not np.any(u == g.dst)  # no self loops
    pairs = set(zip(u.tolist(), g.dst.tolist()))
    assert len(pairs) == g.m  # no duplicates
    assert g.len.min() >= 1 and g.len.max() <= 100
    assert _symmetric(g)


def test_query_bounds_are_integer_floors():
    g = W.grid()
    q = W.random_queries(g, 40, (13, 10), seed=4)
    reach = q.mindelay >= 0
    assert np.all(q.bound[reach] == (q.mindelay[reach] * 13 // 10).astype(np.float32))
    assert np.all(q.src != q.dst)


def test_product_ar_matches_reference_golden():
    with open(GOLDEN) as f:
        gold = json.load(f)
    L = _lib.lib()
    fn = L.vine_ar_burg
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    for case in gold["ar"]:
        x = np.asarray(case["series"], np.float64)
        c = np.zeros(3)
        assert fn(x.ctypes.data, x.shape[0], 3, c.ctypes.data) == 0
        assert c.tolist() == case["burg"]


def test_product_dop_rule_matches_reference_golden():
    with open(GOLDEN) as f:
        gold = json.load(f)
    L = _lib.lib()
    fn = L.vine_dop_sequence
    fn.restype = ctypes.c_int64
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
    for case in gold["dop"]:
        sizes = np.asarray(case["sizes"], np.int64)
        out = np.zeros(sizes.shape[0], np.int64)
        assert fn(sizes.ctypes.data, sizes.shape[0], case["max_labels"], out.ctypes.data) == sizes.shape[0]
        assert out.tolist() == case["dop"]


def test_cpp_facade_argument_check():
    """The C++ facade (include/vine/cspp.hpp) builds with a plain host compiler
    and rejects mismatched task arrays before touching the device."""
    import subprocess
    exe = os.path.join(ROOT, "vine_amd", "vine_facade_check")
    assert os.path.exists(exe), "built by vine_amd.build"
    r = subprocess.run([exe, "args"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "args ok" in r.stdout


def test_window_labels_are_batch_aware():
    """Window labelling for the learned scheduler: a query that sets the
    batch's round count gets the wide window, the rest the least-work one."""
    from vine_amd.learn import label_windows
    E = np.array([[100, 400], [100, 400], [1000, 4000]], float)  # labels expanded per window
    R = np.array([[10, 5], [10, 5], [500, 100]], float)          # rounds per window
    pick, budget, cost = label_windows(E, R, round_cost=50.0)
    assert list(pick) == [0, 0, 1] and budget == 100
    assert cost == 100 + 100 + 4000 + 50.0 * 100


def _landmarks(g):
    L = _lib.lib()
    fn = L.vine_landmark_delays
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    row = np.ascontiguousarray(g.row_ptr, np.int32)
    dst = np.ascontiguousarray(g.dst, np.int32)
    dly = np.ascontiguousarray(g.delay, np.float32)
    out = np.zeros((g.n, 4), np.float32)
    assert fn(g.n, row.ctypes.data, dst.ctypes.data, dly.ctypes.data, out.ctypes.data) == 0
    return out


@pytest.mark.parametrize("kind", ["knn", "rmat"])
def test_landmark_delays_are_shortest_delays_and_bound_the_rest(kind):
    """The landmark delays the forward bound search prunes with
    (vine_landmark_delays, kept by vine_graph_create): each column is scipy's
    Dijkstra from a landmark (the vertex at delay 0), the landmarks are
    distinct, and the bound the search takes on the delay from v to t,
    max_j d(L_j, t) - d(L_j, v) (and, on the symmetric k-NN graph, also
    d(L_j, v) - d(L_j, t)), never exceeds scipy's delay from v to t.  A bound
    above the true delay would prune vertices of feasible paths."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    g = W.knn(20000, seed=5) if kind == "knn" else W.rmat(scale=12, edge_factor=6, seed=2)
    lm = _landmarks(g)
    A = csr_matrix((g.delay.astype(np.float64), g.dst, g.row_ptr), shape=(g.n, g.n))
    seeds = []
    for j in range(4):
        zero = np.nonzero(lm[:, j] == 0.0)[0]
        assert len(zero) >= 1
        s = int(zero[0])
        seeds.append(s)
        d = dijkstra(A, indices=s)
        want = np.where(np.isfinite(d), d, -1.0)
        assert np.array_equal(lm[:, j].astype(np.float64), want), j
    assert len(set(seeds)) == 4
    twin = kind == "knn"
    rng = np.random.default_rng(0)
    targets = rng.choice(g.n, 12, replace=False)
    to_t = dijkstra(A.T.tocsr(), indices=targets)  # delay from every v to each target
    for k, t in enumerate(targets):
        a, b = lm[t][None, :].astype(np.float64), lm.astype(np.float64)
        ok = (a >= 0) & (b >= 0)
        h = np.where(ok, a - b, 0.0)
        if twin:
            h = np.maximum(h, np.where(ok, b - a, 0.0))
        h = np.maximum(h.max(axis=1), 0.0)
        reach = np.isfinite(to_t[k])
        assert np.all(h[reach] <= to_t[k][reach]), (kind, t)


def test_locality_order_is_a_permutation_that_packs_neighbours():
    """The internal vertex order (vine_locality_order, used by vine_graph_create):
    a permutation, each component contiguous, and on the road-like k-NN graph
    arcs span far fewer ids than under the generator's random ids."""
    L = _lib.lib()
    fn = L.vine_locality_order
    fn.restype = ctypes.c_int32
    fn.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    g = W.knn(20000, seed=3)
    row = np.ascontiguousarray(g.row_ptr, np.int32)
    dst = np.ascontiguousarray(g.dst, np.int32)
    order = np.zeros(g.n, np.int32)
    assert fn(g.n, row.ctypes.data, dst.ctypes.data, order.ctypes.data) == 0
    assert np.array_equal(np.sort(order), np.arange(g.n))
    new = np.empty(g.n, np.int64)
    new[order] = np.arange(g.n)
    src = np.repeat(np.arange(g.n), np.diff(row))
    before = np.abs(src - dst).mean()
    after = np.abs(new[src] - new[dst]).mean()
    assert after < before / 20, (before, after)
    # vertices of one component are contiguous in the new order
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    ncomp, lab = connected_components(csr_matrix((np.ones(dst.shape[0]), dst, row), shape=(g.n, g.n)),
                                      directed=True, connection="weak")
    comp_in_order = lab[order]
    changes = np.count_nonzero(comp_in_order[1:] != comp_in_order[:-1])
    assert changes == ncomp - 1
    # a graph with isolated vertices and a directed chain
    row2 = np.array([0, 1, 1, 2, 2], np.int32)
    dst2 = np.array([1, 3], np.int32)
    o2 = np.zeros(4, np.int32)
    assert fn(4, row2.ctypes.data, dst2.ctypes.data, o2.ctypes.data) == 0
    assert sorted(o2.tolist()) == [0, 1, 2, 3]



def test_autotune_cache_round_trip(tmp_path):
    """The autotuner's cache key follows the graph, batch size and policy, and a
    cached choice is returned unchanged (no GPU: the search itself is not run)."""
    from vine_amd import autotune as AT

    g = W.grid(8, 8, seed=1)
    h = W.grid(8, 8, seed=2)
    k = AT.graph_key(g, 64, S.DELTA_MLP)
    assert k == AT.graph_key(W.grid(8, 8, seed=1), 64, S.DELTA_MLP)
    assert k != AT.graph_key(h, 64, S.DELTA_MLP) and k != AT.graph_key(g, 32, S.DELTA_MLP)
    assert k != AT.graph_key(g, 64, S.DELTA_FIXED)
    cache = tmp_path / "tune.json"
    assert AT.cached_options(g, cache=str(cache), max_batch=64, delta_policy=S.DELTA_MLP) is None
    cache.write_text(json.dumps({k: {"options": {"concurrency": 2, "block_size": 64, "grid_blocks": 768}}}))
    got = AT.cached_options(g, cache=str(cache), max_batch=64, delta_policy=S.DELTA_MLP)
    assert got == {"concurrency": 2, "block_size": 64, "grid_blocks": 768}
    assert AT.cached_options(h, cache=str(cache), max_batch=64, delta_policy=S.DELTA_MLP) is None
