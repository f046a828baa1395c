"""Writes tests/golden/answers_<workload>.npz: the exact answer of every query
of the full-size workloads, computed on the CPU by the oracle, for the GPU
parity tests to compare against bit for bit.

Workloads (vine_amd.workload presets, seed 0):
  ny_batch   4,096 queries of the batched NY-scale workload
  ny_1.1x    1,000 random pairs on the NY-scale graph, bound floor(1.1 x min delay)
  ny_1.5x    the same pairs, bound floor(1.5 x min delay)
  ne_1.2x    1,000 random pairs on the NE-scale graph (seed 1), floor(1.2 x)
  rmat17     500 pairs of the R-MAT scale-17 graph's largest component, floor(1.3 x)

Every answer comes from oracle_cspp_astar (label-setting, finishes every query).
A second pass (--cross-check) solves each query again with oracle_cspp (the
round-synchronous restatement of the reference's RUN_CSPP_CPU, with its
lower-bound pruning) under a CPU-time limit per query and rewrites the file:
`restated` = 1 marks the queries it finished (0 = stopped at the limit, 2 = not
run yet), and the pass stops if any of them disagrees with the label-setting
answer.  The arrays also hold the inputs (src, dst, bound) and a CRC of the
graph, so a test can confirm it solves the same workload, and the
label-setting search's work counters (labels expanded X, arcs relaxed E,
labels pushed I: the least work an exact label search does for the query).

Run: python tests/golden/make_full_answers.py [workload ...] [--threads N]
     python tests/golden/make_full_answers.py --cross-check [workload ...] [--limit SECONDS]
"""
import argparse
import os
import sys
import time
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from vine_amd import workload as W  # noqa: E402

WORKLOADS = ("ny_batch", "ny_1.1x", "ny_1.5x", "ne_1.2x", "rmat17")


def load(name):
    """(graph, queries) of a workload, exactly as the GPU tests build them."""
    if name == "ny_batch":
        return W.preset("ny_batch")
    if name in ("ny_1.1x", "ny_1.5x"):
        g, q11, q15 = W.preset("ny")
        return g, (q11 if name == "ny_1.1x" else q15)
    if name == "ne_1.2x":
        return W.preset("ne")
    if name == "rmat17":
        return W.preset("rmat17")
    raise ValueError(name)


def graph_crc(g):
    c = zlib.crc32(np.ascontiguousarray(g.row_ptr).tobytes())
    for a in (g.dst, g.len, g.delay):
        c = zlib.crc32(np.ascontiguousarray(a).tobytes(), c)
    return c


def path(name):
    return os.path.join(HERE, "answers_%s.npz" % name)


def solve_all(name, threads):
    t0 = time.time()
    g, q = load(name)
    nq = len(q)
    print("%s: n=%d m=%d, %d queries (%.1fs)" % (name, g.n, g.m, nq, time.time() - t0), flush=True)

    def astar(i):
        r = O.cspp_astar(g, q.src[i], q.dst[i], q.bound[i])
        if r.feasible:
            assert r.path is not None and r.path[0] == q.src[i] and r.path[-1] == q.dst[i]
        return r

    t0 = time.time()
    with ThreadPoolExecutor(threads) as ex:
        ra = list(ex.map(astar, range(nq)))
    print("  label-setting oracle: %.1fs, %d feasible" % (time.time() - t0, sum(r.feasible for r in ra)),
          flush=True)
    np.savez_compressed(
        path(name),
        n=np.int64(g.n), m=np.int64(g.m), graph_crc=np.int64(graph_crc(g)),
        src=q.src, dst=q.dst, bound=q.bound,
        feasible=np.array([r.feasible for r in ra], np.uint8),
        cost=np.array([r.cost for r in ra], np.float32),
        delay=np.array([r.delay for r in ra], np.float32),
        hops=np.array([0 if r.path is None else max(len(r.path) - 1, 0) for r in ra], np.int32),
        X=np.array([r.counters["expanded"] for r in ra], np.int64),
        E=np.array([r.counters["relaxed"] for r in ra], np.int64),
        I=np.array([r.counters["inserted"] for r in ra], np.int64),
        restated=np.full(nq, 2, np.uint8), restated_limit_s=np.float64(0))
    print("  wrote %s" % os.path.relpath(path(name), ROOT), flush=True)


def cross_check(name, threads, limit):
    fx = dict(np.load(path(name)))
    t0 = time.time()
    g, q = load(name)
    assert np.array_equal(fx["src"], q.src) and np.array_equal(fx["bound"], q.bound)
    assert int(fx["graph_crc"]) == graph_crc(g)
    nq = len(q)

    def restated(i):
        try:
            return O.cspp(g, q.src[i], q.dst[i], q.bound[i], prune_target=2, time_limit=limit)
        except O.TimeLimit:
            return None

    with ThreadPoolExecutor(threads) as ex:
        rr = list(ex.map(restated, range(nq)))
    done = np.array([r is not None for r in rr], np.uint8)
    for i, r in enumerate(rr):
        if r is None:
            continue
        f = bool(fx["feasible"][i])
        same = r.feasible == f and (not f or (np.float32(r.cost) == fx["cost"][i]
                                              and np.float32(r.delay) == fx["delay"][i]))
        if not same:
            raise SystemExit("%s query %d: restated oracle (%s, %r, %r) != label-setting (%s, %r, %r)"
                             % (name, i, r.feasible, r.cost, r.delay, f, fx["cost"][i], fx["delay"][i]))
    fx["restated"] = done
    fx["restated_limit_s"] = np.float64(limit)
    np.savez_compressed(path(name), **fx)
    print("%s: restated oracle %.1fs, %d/%d finished within %.0f s CPU each, all agree" %
          (name, time.time() - t0, int(done.sum()), nq, limit), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--cross-check", action="store_true")
    ap.add_argument("--limit", type=float, default=20.0, help="CPU seconds per query for oracle_cspp")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    for name in args.workloads:
        if args.cross_check:
            cross_check(name, args.threads, args.limit)
        else:
            solve_all(name, args.threads)


if __name__ == "__main__":
    main()
