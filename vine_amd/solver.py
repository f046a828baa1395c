"""Python binding of the CSP solver C ABI (include/vine.h).

`Solver` is the handle-based interface; `cspp` mirrors the reference facade
cspp<Label_t, Edge_t> (algorithm/cspp.hpp:11-69): constructed from a graph, the
number of tasks, the source labels, the sinks and the per-task constraints, with
solve() returning 0 and leaving results on the object.  Every call goes to the
HIP path in libvine.so; there is no CPU fallback.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import lib

PRUNE_NONE, PRUNE_TARGET, PRUNE_BOUNDS = 0, 1, 2
DELTA_INF, DELTA_FIXED, DELTA_AR, DELTA_MLP, DELTA_DOP = 0, 1, 2, 3, 4
ENGINE_AUTO, ENGINE_ROUNDS, ENGINE_WORKGROUP = 0, 1, 2

VINE_OK, VINE_E_ARG, VINE_E_OOM, VINE_E_HIP, VINE_E_OVERFLOW, VINE_E_RANGE = 0, -1, -2, -3, -4, -5
_STATUS = {VINE_E_ARG: "VINE_E_ARG", VINE_E_OOM: "VINE_E_OOM", VINE_E_HIP: "VINE_E_HIP",
           VINE_E_OVERFLOW: "VINE_E_OVERFLOW", VINE_E_RANGE: "VINE_E_RANGE"}


class VineError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (_STATUS.get(status, status), msg))
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [
        ("max_batch", ctypes.c_int32),
        ("prune", ctypes.c_int32),
        ("delta_policy", ctypes.c_int32),
        ("delta_fixed", ctypes.c_float),
        ("use_graph", ctypes.c_int32),
        ("rounds_per_sync", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("grid_blocks", ctypes.c_int32),
        ("hash_slots", ctypes.c_int32),
        ("label_capacity", ctypes.c_int64),
        ("cand_capacity", ctypes.c_int64),
        ("pool_capacity", ctypes.c_int64),
        ("frontier_capacity", ctypes.c_int64),
        ("ar_max_labels", ctypes.c_double),
        ("record_trace", ctypes.c_int32),
        ("time_kernels", ctypes.c_int32),
        ("bound_delta", ctypes.c_float),
        ("bound_ball", ctypes.c_int32),
        ("engine", ctypes.c_int32),
        ("labels_per_wave", ctypes.c_int32),
        ("concurrency", ctypes.c_int32),
        ("near_windows", ctypes.c_float),
        ("lagrange_bound", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int64),
        ("sssp_rounds", ctypes.c_int64),
        ("expanded", ctypes.c_int64),
        ("relaxed", ctypes.c_int64),
        ("candidates", ctypes.c_int64),
        ("inserted", ctypes.c_int64),
        ("sssp_relaxed", ctypes.c_int64),
        ("max_frontier", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("ms_total", ctypes.c_double),
        ("ms_kernels", ctypes.c_double),
        ("ms_expand", ctypes.c_double),
        ("ms_merge", ctypes.c_double),
        ("ms_sssp", ctypes.c_double),
        ("expand_launches", ctypes.c_int64),
        ("merge_launches", ctypes.c_int64),
        ("sssp_launches", ctypes.c_int64),
        ("touched", ctypes.c_int64),
        ("merge_scanned", ctypes.c_int64),
        ("merge_chain_max", ctypes.c_int64),
        ("far_pulled", ctypes.c_int64),
        ("ms_pull", ctypes.c_double),
        ("near_carried", ctypes.c_int64),
        ("far_appended", ctypes.c_int64),
        ("scanned", ctypes.c_int64),
        ("timed_rounds", ctypes.c_int64),
        ("far_units", ctypes.c_int64),
        ("cand_chunks", ctypes.c_int64),
        ("hash_dropped", ctypes.c_int64),
        ("hash_overflow", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


EXPORTS = (
    "vine_options_default", "vine_last_error", "vine_device_count", "vine_graph_create",
    "vine_graph_destroy", "vine_solver_create", "vine_solver_destroy", "vine_solve",
    "vine_solve_device", "vine_solver_latency", "vine_solver_trace", "vine_solver_set_mlp",
    "vine_solver_pareto", "vine_solver_query_stats", "vine_family_solver_create", "vine_family_solver_destroy",
    "vine_family_solve", "vine_family_front", "vine_solver_paths", "vine_solve_device_static",
)

_declared = False


def _L():
    global _declared
    L = lib()
    if not _declared:
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        P = ctypes.POINTER
        sig = {
            "vine_options_default": (None, [P(Options)]),
            "vine_last_error": (ctypes.c_char_p, []),
            "vine_device_count": (i32, []),
            "vine_graph_create": (i32, [i32, i64, vp, vp, vp, vp, P(vp)]),
            "vine_graph_destroy": (i32, [vp]),
            "vine_solver_create": (i32, [vp, P(Options), P(vp)]),
            "vine_solver_destroy": (i32, [vp]),
            "vine_solve": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, i64, P(Stats)]),
            "vine_solve_device": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, P(Stats)]),
            "vine_solver_latency": (i32, [vp, i32, vp]),
            "vine_solver_trace": (i64, [vp, vp, i64]),
            "vine_solver_query_stats": (i32, [vp, i32, vp, vp]),
            "vine_solver_set_mlp": (i32, [vp, vp, i64]),
            "vine_solver_pareto": (i64, [vp, i32, i32, vp, i64]),
            "vine_solver_paths": (i32, [vp, i32, vp, vp, i64]),
            "vine_solve_device_static": (i32, [vp, P(Options), vp, i64, i32, vp, vp, vp, vp, vp, vp, vp, P(Stats)]),
            "vine_family_solver_create": (i32, [vp, i32, i32, i64, i32, ctypes.c_float, P(vp)]),
            "vine_family_solver_destroy": (i32, [vp]),
            "vine_family_solve": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, i64, vp]),
            "vine_family_front": (i64, [vp, i32, i32, vp, i64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _declared = True
    return L


def _check(rc):
    if rc != VINE_OK:
        raise VineError(rc, _L().vine_last_error().decode(errors="replace"))


def default_options(**kw):
    o = Options()
    _L().vine_options_default(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError("unknown option %r" % k)
        setattr(o, k, v)
    return o


def device_count():
    return int(_L().vine_device_count())


class DeviceGraph:
    """Device-resident CSR (counterpart of Graph::make_gpu_graph, include/graph.hpp:169-195)."""

    def __init__(self, g):
        self.n, self.m = g.n, g.m
        self._keep = (np.ascontiguousarray(g.row_ptr, np.int32), np.ascontiguousarray(g.dst, np.int32),
                      np.ascontiguousarray(g.len, np.float32), np.ascontiguousarray(g.delay, np.float32))
        r, d, l, dl = self._keep
        h = ctypes.c_void_p()
        _check(_L().vine_graph_create(self.n, self.m, r.ctypes.data, d.ctypes.data, l.ctypes.data,
                                      dl.ctypes.data, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _L().vine_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Answer:
    cost: np.ndarray
    delay: np.ndarray
    feasible: np.ndarray
    path_off: np.ndarray
    path: np.ndarray
    stats: dict

    def path_of(self, i):
        return self.path[self.path_off[i]:self.path_off[i + 1]]


class Solver:
    def __init__(self, graph, **options):
        self.graph = graph if isinstance(graph, DeviceGraph) else DeviceGraph(graph)
        self.options = default_options(**options)
        # path vertices per query seen so far: sizes the next path buffer so
        # that a solve is not repeated for want of room (VINE_E_OVERFLOW)
        self._path_per_query = 0.0
        h = ctypes.c_void_p()
        _check(_L().vine_solver_create(self.graph.h, ctypes.byref(self.options), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _L().vine_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve(self, src, dst, bound, paths=True, path_cap=None):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        bound = np.ascontiguousarray(bound, np.float32)
        nq = int(src.shape[0])
        if dst.shape[0] != nq or bound.shape[0] != nq:
            raise ValueError("src, dst and bound must have the same length")
        cost = np.zeros(nq, np.float32)
        delay = np.zeros(nq, np.float32)
        feas = np.zeros(nq, np.int32)
        off = np.zeros(nq + 1, np.int64)
        if path_cap is None:
            path_cap = max(1, nq * max(256, int(1.25 * self._path_per_query) + 1))
        while True:
            path = np.zeros(path_cap if paths else 1, np.int32)
            st = Stats()
            rc = _L().vine_solve(self.h, nq, src.ctypes.data, dst.ctypes.data, bound.ctypes.data,
                                 cost.ctypes.data, delay.ctypes.data, feas.ctypes.data, off.ctypes.data,
                                 path.ctypes.data if paths else None, path_cap if paths else 0,
                                 ctypes.byref(st))
            if rc == VINE_E_OVERFLOW and paths and off[-1] > path_cap:
                path_cap = int(off[-1])
                if nq <= self.options.max_batch:
                    # one batch: its paths are still on the device, read them again
                    path = np.zeros(path_cap, np.int32)
                    _check(_L().vine_solver_paths(self.h, nq, off.ctypes.data, path.ctypes.data, path_cap))
                    break
                continue
            _check(rc)
            break
        if paths and nq:
            self._path_per_query = max(self._path_per_query, float(off[-1]) / nq)
        return Answer(cost, delay, feas.astype(bool), off, path[:off[-1]] if paths else None, st.as_dict())

    def solve_device(self, d_src, d_dst, d_bound, d_cost, d_delay, d_feas, nq):
        """All arguments are device addresses (ints); returns stats."""
        st = Stats()
        _check(_L().vine_solve_device(self.h, int(nq), d_src, d_dst, d_bound, d_cost, d_delay, d_feas,
                                      ctypes.byref(st)))
        return st.as_dict()

    def latency_us(self, nq):
        out = np.zeros(nq, np.float64)
        _check(_L().vine_solver_latency(self.h, nq, out.ctypes.data))
        return out

    def trace(self):
        """Scheduler feature rows of the last solve: columns (query, round, x0..x7)."""
        n = _L().vine_solver_trace(self.h, None, 0)
        rows = np.zeros((max(n, 1), 10), np.float32)
        k = _L().vine_solver_trace(self.h, rows.ctypes.data, n)
        return rows[:k]

    def query_stats(self, nq):
        """(labels expanded, active rounds) per query of the last solve."""
        ex = np.zeros(nq, np.int64)
        rd = np.zeros(nq, np.int32)
        _check(_L().vine_solver_query_stats(self.h, nq, ex.ctypes.data, rd.ctypes.data))
        return ex, rd

    def pareto(self, q, v, cap=1 << 16):
        out = np.zeros(2 * cap, np.float32)
        n = _L().vine_solver_pareto(self.h, int(q), int(v), out.ctypes.data, cap)
        if n < 0:
            _check(int(n))
        k = min(n, cap)
        pairs = out[:2 * k].reshape(k, 2)
        order = np.lexsort((pairs[:, 1], pairs[:, 0]))
        return pairs[order]

    def set_mlp(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(_L().vine_solver_set_mlp(self.h, p.ctypes.data, p.shape[0]))


# label families (vine_family in include/vine.h; the reference's CSP values)
FAMILY_CSP, FAMILY_HOP, FAMILY_HOP_MAXCOST, FAMILY_2OSP, FAMILY_SOCIAL, FAMILY_QOS = 1, 2, 3, 4, 5, 6
# the reference's constructor defaults for hop_L and maxcost_L (include/constraint.hpp:31, 61)
DEFAULT_HOP_L, DEFAULT_MAXCOST_L = 100, 10000.0


@dataclass
class FamilyAnswer:
    label: np.ndarray  # float32[nq, 4]: d, delay, hop, maxcost of the best label at t (+inf if infeasible)
    feasible: np.ndarray
    path_off: np.ndarray
    path: np.ndarray
    counters: dict

    def path_of(self, i):
        return self.path[self.path_off[i]:self.path_off[i + 1]]


class FamilySolver:
    """Generic label-correcting search for one label family (vine_family_solve)."""

    def __init__(self, graph, family, max_batch=64, label_capacity=0, prune=0, window=0.0):
        """prune = 0: the reference's one-to-all search (complete Pareto sets);
        1: drop labels that cannot beat t's best answer (same answers).
        window > 0: key-ordered rounds (length + bound to t within window of the
        query's smallest); 0: every frontier label every round."""
        self.graph = graph if isinstance(graph, DeviceGraph) else DeviceGraph(graph)
        h = ctypes.c_void_p()
        _check(_L().vine_family_solver_create(self.graph.h, int(family), int(max_batch), int(label_capacity),
                                              int(prune), float(window), ctypes.byref(h)))
        self.h = h
        self.family = family

    def close(self):
        if getattr(self, "h", None):
            _L().vine_family_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def solve(self, src, dst, delay_L, hop_L=DEFAULT_HOP_L, maxcost_L=DEFAULT_MAXCOST_L, paths=True):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        nq = int(src.shape[0])
        lim = np.zeros((nq, 3), np.float32)
        lim[:, 0] = delay_L
        lim[:, 1] = hop_L
        lim[:, 2] = maxcost_L
        out = np.zeros((nq, 4), np.float32)
        feas = np.zeros(nq, np.int32)
        off = np.zeros(nq + 1, np.int64)
        cnt = np.zeros(4, np.int64)
        path_cap = max(1, nq * 256)  # retried at the reported size if short
        while True:
            path = np.zeros(path_cap if paths else 1, np.int32)
            rc = _L().vine_family_solve(self.h, nq, src.ctypes.data, dst.ctypes.data, lim.ctypes.data,
                                        out.ctypes.data, feas.ctypes.data, off.ctypes.data,
                                        path.ctypes.data if paths else None, path_cap if paths else 0,
                                        cnt.ctypes.data)
            if rc == VINE_E_OVERFLOW and paths and off[-1] > path_cap:
                path_cap = int(off[-1])
                continue
            _check(rc)
            break
        return FamilyAnswer(out, feas.astype(bool), off, path[:off[-1]] if paths else None,
                            dict(zip(("rounds", "expanded", "relaxed", "inserted"), (int(x) for x in cnt))))

    def front(self, q, v, cap=1 << 12):
        """Pareto set of vertex v for query q of the last batch, (k, 4) float32."""
        out = np.zeros(4 * cap, np.float32)
        n = _L().vine_family_front(self.h, int(q), int(v), out.ctypes.data, cap)
        if n < 0:
            _check(int(n))
        return out[:4 * min(n, cap)].reshape(-1, 4).copy()


class c_label:
    """Source label (include/label.hpp:183-189): vertex id with both sums zero."""

    def __init__(self, vid, d=0.0, delay=0.0):
        self.vid, self.d, self.delay = int(vid), float(d), float(delay)


class c_constraint:
    """Delay bound, inclusive (include/constraint.hpp:81-105)."""

    def __init__(self, delay_L):
        self.delay_L = float(delay_L)

    def pass_constraints_check(self, label):
        return not (label.delay > self.delay_L)


class cspp:
    """Mirror of cspp<c_label, c_edge> (algorithm/cspp.hpp:11-69).

    cspp(g, ntask, source_state, sink_l, C_l).solve() runs every task on the
    GPU and leaves cost_l, delay_l, feasible_l, path_l and size_of_result_l on
    the object (the reference allocates path_l but never fills it,
    algorithm/cspp.hpp:36-46)."""

    def __init__(self, g, ntask, source_state, sink_l, C_l, **options):
        if ntask != len(source_state) or ntask != len(sink_l) or ntask != len(C_l):
            raise ValueError("ntask must match the lengths of source_state, sink_l and C_l")
        self.g = g
        self.ntask = ntask
        self.init_state = list(source_state)
        self.sink_l = list(sink_l)
        self.C_l = list(C_l)
        self.options = options
        self.path_l = []
        self.size_of_result_l = []

    def solve(self):
        src = [l.vid if isinstance(l, c_label) else int(l) for l in self.init_state]
        bnd = [c.delay_L if isinstance(c, c_constraint) else float(c) for c in self.C_l]
        with Solver(self.g, **self.options) as s:
            a = s.solve(src, self.sink_l, bnd)
        self.cost_l = a.cost
        self.delay_l = a.delay
        self.feasible_l = a.feasible
        self.path_l = [a.path_of(i) for i in range(self.ntask)]
        self.size_of_result_l = [len(p) for p in self.path_l]
        self.stats = a.stats
        return 0
