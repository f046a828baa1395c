"""Seeded synthetic workloads (graphs and query sets) for the CSP solver.

Thin Python layer over the generators in libvine.so (vine_amd/csrc/workload.cpp).
Graphs are returned as CSR numpy arrays in adjacency insertion order, the order
the reference's Graph::adj_to_csr produces (include/graph.hpp:136-167).

Standard workloads:
  grid64    64x64 4-neighbour grid, 10% of edges deleted (seed 0), len U{1..1000},
            delay U{1..100}; 100 random pairs; bound floor(1.2 * min delay).
  ny        264,346 points, 3-NN symmetrised, len = round(dist * 1e5),
            delay = round(len * U[0.5, 2]); random pairs, bounds 1.1x / 1.5x.
  ne        same generator with 1,524,453 points; bound 1.2x.
  rmat17    R-MAT scale 17, edge factor 8, (0.57, 0.19, 0.19, 0.05); len, delay
            U{1..100}; pairs from the largest component; bound 1.3x.
  ny_batch  ny graph, 4,096 queries with hop distance log-uniform in [10, 2000]
            and bound factor U[1.05, 2.0].
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import lib


@dataclass
class Graph:
    row_ptr: np.ndarray  # int32[n+1]
    dst: np.ndarray  # int32[m]
    len: np.ndarray  # float32[m]
    delay: np.ndarray  # float32[m]
    name: str = ""
    seed: int = 0

    @property
    def n(self):
        return int(self.row_ptr.shape[0] - 1)

    @property
    def m(self):
        return int(self.dst.shape[0])


@dataclass
class Queries:
    src: np.ndarray  # int32[q]
    dst: np.ndarray  # int32[q]
    bound: np.ndarray  # float32[q]
    mindelay: np.ndarray  # int64[q], -1 if unreachable (None for queries read from a file)
    hops: np.ndarray = field(default=None)  # int32[q] (hop-distance queries only)
    aux: np.ndarray = field(default=None)  # int32[q, 2] (read from a file: (lower, upper) or (hop, cost))

    def __len__(self):
        return int(self.src.shape[0])

    def subset(self, idx):
        idx = np.asarray(idx)
        take = lambda a: None if a is None else a[idx].copy()
        return Queries(self.src[idx].copy(), self.dst[idx].copy(), self.bound[idx].copy(),
                       take(self.mindelay), take(self.hops), take(self.aux))


def _take(h, name="", seed=0):
    L = lib()
    n = L.vine_host_graph_n(h)
    m = L.vine_host_graph_m(h)

    def arr(fn, ct, count, dt):
        p = fn(h)
        if count == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(count,)).astype(dt, copy=True)

    g = Graph(arr(L.vine_host_graph_row_ptr, ctypes.c_int32, n + 1, np.int32),
              arr(L.vine_host_graph_dst, ctypes.c_int32, m, np.int32),
              arr(L.vine_host_graph_len, ctypes.c_float, m, np.float32),
              arr(L.vine_host_graph_delay, ctypes.c_float, m, np.float32), name, seed)
    L.vine_host_graph_free(h)
    return g


def _check(rc, what):
    if rc != 0:
        raise ValueError("%s failed (status %d)" % (what, rc))


def grid(rows=64, cols=64, delete_permille=100, len_range=(1, 1000), delay_range=(1, 100), seed=0):
    h = ctypes.c_void_p()
    _check(lib().vine_gen_grid(rows, cols, delete_permille, len_range[0], len_range[1],
                               delay_range[0], delay_range[1], seed, ctypes.byref(h)), "vine_gen_grid")
    return _take(h, "grid%dx%d" % (rows, cols), seed)


def knn(n, k=3, len_scale=1e5, seed=0):
    h = ctypes.c_void_p()
    _check(lib().vine_gen_knn(n, k, len_scale, seed, ctypes.byref(h)), "vine_gen_knn")
    return _take(h, "knn%d_k%d" % (n, k), seed)


def rmat(scale=17, edge_factor=8, a=0.57, b=0.19, c=0.19, lo=1, hi=100, seed=0):
    h = ctypes.c_void_p()
    _check(lib().vine_gen_rmat(scale, edge_factor, a, b, c, lo, hi, seed, ctypes.byref(h)), "vine_gen_rmat")
    return _take(h, "rmat%d" % scale, seed)


def from_csr(row_ptr, dst, len_, delay, name="custom"):
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    len_ = np.ascontiguousarray(len_, np.float32)
    delay = np.ascontiguousarray(delay, np.float32)
    h = ctypes.c_void_p()
    _check(lib().vine_host_graph_from_csr(row_ptr.shape[0] - 1, dst.shape[0], row_ptr.ctypes.data,
                                          dst.ctypes.data, len_.ctypes.data, delay.ctypes.data,
                                          ctypes.byref(h)), "vine_host_graph_from_csr")
    return _take(h, name)


FILE_SNAP = 1  # SNAP edge list (the reference's graph_type 1), undirected
FILE_COLA = 2  # COLA road network (graph_type 2), directed arcs


def _read_error(what, rc):
    msg = lib().vine_last_error()
    raise ValueError("%s failed (status %d): %s" % (what, rc, msg.decode() if msg else ""))


def read_graph(path, fmt):
    """Graph file in the reference's formats (read_graph, include/io.hpp:22-91):
    FILE_COLA ("u: v w c ... -1" lines, arcs as listed) or FILE_SNAP (undirected
    rows, the reference's "N?M?" layout or the shipped MatrixMarket files)."""
    h = ctypes.c_void_p()
    rc = lib().vine_read_graph(str(path).encode(), fmt, ctypes.byref(h))
    if rc != 0:
        _read_error("vine_read_graph", rc)
    return _take(h, os.path.basename(str(path)))


def read_queries(path, fmt):
    """Query file (task::read_query / load_query, include/task.hpp:27-76):
    FILE_COLA rows "u v lower upper limit" (bound = limit, aux = (lower, upper))
    or FILE_SNAP CSV rows "id,src,snk,delay,hop,cost" after a header line
    (bound = delay, aux = (hop, cost))."""
    path = str(path).encode()
    nq = lib().vine_read_queries(path, fmt, 0, None, None, None, None)
    if nq < 0:
        _read_error("vine_read_queries", nq)
    src = np.zeros(nq, np.int32)
    dst = np.zeros(nq, np.int32)
    bound = np.zeros(nq, np.float32)
    aux = np.zeros((nq, 2), np.int32)
    got = lib().vine_read_queries(path, fmt, nq, src.ctypes.data, dst.ctypes.data, bound.ctypes.data,
                                  aux.ctypes.data)
    if got != nq:
        _read_error("vine_read_queries", got)
    return Queries(src, dst, bound, None, aux=aux)


class _HostHandle:
    """Temporary native handle for a Graph (for the query generators)."""

    def __init__(self, g):
        self.h = ctypes.c_void_p()
        _check(lib().vine_host_graph_from_csr(g.n, g.m, g.row_ptr.ctypes.data, g.dst.ctypes.data,
                                              g.len.ctypes.data, g.delay.ctypes.data,
                                              ctypes.byref(self.h)), "vine_host_graph_from_csr")

    def __enter__(self):
        return self.h

    def __exit__(self, *a):
        lib().vine_host_graph_free(self.h)


def random_queries(g, nq, factor=(12, 10), lcc_only=False, seed=0, nthreads=0):
    src = np.zeros(nq, np.int32)
    dst = np.zeros(nq, np.int32)
    bound = np.zeros(nq, np.float32)
    md = np.zeros(nq, np.int64)
    with _HostHandle(g) as h:
        _check(lib().vine_gen_queries_random(h, nq, 1 if lcc_only else 0, factor[0], factor[1], seed,
                                             nthreads, src.ctypes.data, dst.ctypes.data,
                                             bound.ctypes.data, md.ctypes.data), "vine_gen_queries_random")
    return Queries(src, dst, bound, md)


def hop_queries(g, nq, hop_lo=10, hop_hi=2000, f_permille=(1050, 2000), seed=0, nthreads=0):
    src = np.zeros(nq, np.int32)
    dst = np.zeros(nq, np.int32)
    bound = np.zeros(nq, np.float32)
    md = np.zeros(nq, np.int64)
    hops = np.zeros(nq, np.int32)
    with _HostHandle(g) as h:
        _check(lib().vine_gen_queries_hops(h, nq, hop_lo, hop_hi, f_permille[0], f_permille[1], seed,
                                           nthreads, src.ctypes.data, dst.ctypes.data, bound.ctypes.data,
                                           md.ctypes.data, hops.ctypes.data), "vine_gen_queries_hops")
    return Queries(src, dst, bound, md, hops)


def min_delay(g, s, t):
    with _HostHandle(g) as h:
        return int(lib().vine_min_delay(h, s, t))


def preset(name, seed=0):
    """Return (graph, queries...) for a named standard workload."""
    if name == "grid64":
        g = grid(seed=seed)
        return g, random_queries(g, 100, (12, 10), seed=seed)
    if name == "ny":
        g = knn(264346, seed=seed)
        return g, random_queries(g, 1000, (11, 10), seed=seed), random_queries(g, 1000, (15, 10), seed=seed)
    if name == "ne":
        g = knn(1524453, seed=seed)
        return g, random_queries(g, 1000, (12, 10), seed=seed + 1)
    if name == "rmat17":
        g = rmat(seed=seed)
        return g, random_queries(g, 500, (13, 10), lcc_only=True, seed=seed)
    if name == "ny_batch":
        g = knn(264346, seed=seed)
        return g, hop_queries(g, 4096, seed=seed)
    raise ValueError("unknown workload %r" % name)
