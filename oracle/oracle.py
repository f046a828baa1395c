"""Python bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never by the product path.

oracle_cspp restates RUN_CSPP_CPU (algorithm/cspp_cpu.hpp:44-125) with the
additive expand of c_label::expand (include/label.hpp:213-222).
"""
import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_AR = os.path.join(HERE, "_ref", "libar_ref.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_double
        L.oracle_cspp.restype = i32
        L.oracle_cspp.argtypes = [i32, vp, vp, vp, vp, i32, i32, f32, i32, vp, vp, vp, i32, vp, vp, i32,
                                  vp, vp, vp, i32, f64]
        L.oracle_cspp_astar.restype = i32
        L.oracle_cspp_astar.argtypes = [i32, vp, vp, vp, vp, i32, i32, f32, vp, vp, vp, i32, vp, vp, f64]
        L.oracle_dijkstra_lex.restype = i32
        L.oracle_dijkstra_lex.argtypes = [i32, vp, vp, vp, vp, i32, vp, vp]
        L.oracle_resource_dp.restype = i32
        L.oracle_resource_dp.argtypes = [i32, vp, vp, vp, vp, i32, i32, i32, vp, vp]
        L.oracle_bruteforce.restype = i32
        L.oracle_bruteforce.argtypes = [i32, vp, vp, vp, vp, i32, i32, f32, vp, vp]
        L.oracle_ar_burg.restype = i32
        L.oracle_ar_burg.argtypes = [vp, i32, i32, vp]
        L.oracle_ar_lsq.restype = i32
        L.oracle_ar_lsq.argtypes = [vp, i32, i32, vp]
        L.oracle_family.restype = i32
        L.oracle_family.argtypes = [i32, i32, vp, vp, vp, vp, i32, i32, vp, vp, vp, i32, vp, vp, i32, vp, vp,
                                    ctypes.c_int64]
        L.oracle_family_bruteforce.restype = i32
        L.oracle_family_bruteforce.argtypes = [i32, i32, vp, vp, vp, vp, i32, i32, vp, vp, i32]
        _lib = L
    return _lib


@dataclass
class Result:
    feasible: bool
    cost: float
    delay: float
    path: np.ndarray
    pareto: np.ndarray  # (k, 2) float32, sorted by d
    counters: dict
    trace: np.ndarray = None


COUNTER_NAMES = ("expanded", "relaxed", "inserted", "rounds", "max_frontier", "sum_frontier")


def _g(g):
    return (g.n, g.row_ptr.ctypes.data, g.dst.ctypes.data, g.len.ctypes.data, g.delay.ctypes.data)


class TimeLimit(RuntimeError):
    pass


def cspp(g, s, t, bound, prune_target=False, path_cap=None, pareto_cap=4096, trace_cap=0, time_limit=0.0):
    """Solve one query; returns Result.  prune_target: 0/False reference
    one-to-all search, 1 target pruning, 2 lower-bound pruning."""
    path_cap = g.n if path_cap is None else path_cap
    cost = ctypes.c_float()
    dly = ctypes.c_float()
    path = np.zeros(max(path_cap, 1), np.int32)
    plen = ctypes.c_int32()
    par = np.zeros(2 * pareto_cap, np.float32)
    parl = ctypes.c_int32()
    cnt = np.zeros(6, np.int64)
    trace = np.zeros(max(trace_cap, 1), np.int32)
    rc = lib().oracle_cspp(*_g(g), int(s), int(t), float(bound), int(prune_target),
                           ctypes.byref(cost), ctypes.byref(dly), path.ctypes.data, path_cap,
                           ctypes.byref(plen), par.ctypes.data, pareto_cap, ctypes.byref(parl),
                           cnt.ctypes.data, trace.ctypes.data if trace_cap else None, trace_cap,
                           float(time_limit))
    if rc == -4:
        raise TimeLimit("oracle_cspp exceeded %.1f s" % time_limit)
    if rc < 0:
        raise RuntimeError("oracle_cspp failed (%d)" % rc)
    k = min(parl.value, pareto_cap)
    counters = dict(zip(COUNTER_NAMES, (int(x) for x in cnt)))
    tr = trace[:min(counters["rounds"], trace_cap)].copy() if trace_cap else None
    return Result(rc == 1, cost.value, dly.value,
                  path[:plen.value].copy() if plen.value <= path_cap else None,
                  par[:2 * k].reshape(k, 2).copy(), counters, tr)


def cspp_astar(g, s, t, bound, path_cap=None, time_limit=0.0):
    """oracle_cspp_astar: the same answer as cspp() by lexicographic bi-criteria
    A* (label-setting, exact reverse-Dijkstra heuristics).  Returns Result
    (pareto empty; counters expanded / relaxed / inserted = labels pushed)."""
    path_cap = g.n if path_cap is None else path_cap
    cost = ctypes.c_float()
    dly = ctypes.c_float()
    path = np.zeros(max(path_cap, 1), np.int32)
    plen = ctypes.c_int32()
    cnt = np.zeros(3, np.int64)
    rc = lib().oracle_cspp_astar(*_g(g), int(s), int(t), float(bound), ctypes.byref(cost), ctypes.byref(dly),
                                 path.ctypes.data, path_cap, ctypes.byref(plen), cnt.ctypes.data,
                                 float(time_limit))
    if rc == -4:
        raise TimeLimit("oracle_cspp_astar exceeded %.1f s" % time_limit)
    if rc < 0:
        raise RuntimeError("oracle_cspp_astar failed (%d)" % rc)
    counters = dict(zip(("expanded", "relaxed", "inserted"), (int(x) for x in cnt)))
    return Result(rc == 1, cost.value, dly.value,
                  path[:plen.value].copy() if plen.value <= path_cap else None,
                  np.zeros((0, 2), np.float32), counters)


@dataclass
class FamilyResult:
    feasible: bool
    best: np.ndarray  # float32[4]: d, delay, hop, maxcost
    path: np.ndarray
    front: np.ndarray  # (k, 4) float32, t's Pareto set in the answer order
    counters: dict


def family(g, fam, s, t, limits, front_cap=65536, max_labels=1 << 28):
    """oracle_family: the label family `fam` (1..6 as vine_family) for one query;
    limits = (delay_L, hop_L, maxcost_L)."""
    lim = np.asarray(limits, np.float32)
    best = np.zeros(4, np.float32)
    path = np.zeros(g.n, np.int32)
    plen = ctypes.c_int32()
    front = np.zeros(4 * front_cap, np.float32)
    flen = ctypes.c_int32()
    cnt = np.zeros(4, np.int64)
    rc = lib().oracle_family(int(fam), *_g(g), int(s), int(t), lim.ctypes.data, best.ctypes.data, path.ctypes.data,
                             g.n, ctypes.byref(plen), front.ctypes.data, front_cap, ctypes.byref(flen),
                             cnt.ctypes.data, int(max_labels))
    if rc < 0:
        raise RuntimeError("oracle_family failed (%d)" % rc)
    k = min(flen.value, front_cap)
    return FamilyResult(rc == 1, best, path[:plen.value].copy(), front[:4 * k].reshape(k, 4).copy(),
                        dict(zip(("rounds", "expanded", "relaxed", "inserted"), (int(x) for x in cnt))))


def family_bruteforce(g, fam, s, t, limits, cap=100000):
    """t's non-dominated set over simple paths (tiny graphs), answer order."""
    lim = np.asarray(limits, np.float32)
    out = np.zeros(4 * cap, np.float32)
    m = lib().oracle_family_bruteforce(int(fam), *_g(g), int(s), int(t), lim.ctypes.data, out.ctypes.data, cap)
    if m < 0:
        raise RuntimeError("oracle_family_bruteforce failed (%d)" % m)
    return out[:4 * min(m, cap)].reshape(-1, 4).copy()


def dijkstra_lex(g, s, w1="len", w2="delay"):
    a = getattr(g, w1) if isinstance(w1, str) else w1
    b = getattr(g, w2) if isinstance(w2, str) else w2
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    d1 = np.zeros(g.n, np.float64)
    d2 = np.zeros(g.n, np.float64)
    rc = lib().oracle_dijkstra_lex(g.n, g.row_ptr.ctypes.data, g.dst.ctypes.data, a.ctypes.data,
                                   b.ctypes.data, int(s), d1.ctypes.data, d2.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_dijkstra_lex failed")
    return d1, d2


def resource_dp(g, s, t, bound):
    c = ctypes.c_double()
    d = ctypes.c_double()
    rc = lib().oracle_resource_dp(*_g(g), int(s), int(t), int(bound), ctypes.byref(c), ctypes.byref(d))
    if rc < 0:
        raise RuntimeError("oracle_resource_dp failed (%d)" % rc)
    return rc == 1, c.value, d.value


def bruteforce(g, s, t, bound):
    c = ctypes.c_double()
    d = ctypes.c_double()
    rc = lib().oracle_bruteforce(*_g(g), int(s), int(t), float(bound), ctypes.byref(c), ctypes.byref(d))
    if rc < 0:
        raise RuntimeError("oracle_bruteforce failed (%d)" % rc)
    return rc == 1, c.value, d.value


def ar_burg(series, degree=3):
    x = np.ascontiguousarray(series, np.float64)
    c = np.zeros(degree, np.float64)
    if lib().oracle_ar_burg(x.ctypes.data, x.shape[0], degree, c.ctypes.data) != 0:
        raise ValueError("oracle_ar_burg: bad arguments")
    return c


def ar_lsq(series, degree=3):
    x = np.ascontiguousarray(series, np.float64)
    c = np.zeros(degree, np.float64)
    if lib().oracle_ar_lsq(x.ctypes.data, x.shape[0], degree, c.ctypes.data) != 0:
        raise ValueError("oracle_ar_lsq failed")
    return c


def ref_ar(series, degree=3, method=0):
    """The reference's own AR::AutoRegression, compiled from include/AR.cpp into
    oracle/_ref/libar_ref.so (None if that build is absent)."""
    if not os.path.exists(REF_AR):
        return None
    L = ctypes.CDLL(REF_AR)
    fn = getattr(L, "_ZN2AR14AutoRegressionEPdiiS0_i")
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    x = np.ascontiguousarray(series, np.float64).copy()
    c = np.zeros(degree, np.float64)
    fn(x.ctypes.data, x.shape[0], degree, c.ctypes.data, method)
    return c
