"""Launch-shape autotuner for the batched CSP solver.

Times a few solver configurations on the caller's own graph and queries and
keeps the fastest in a JSON cache, keyed by the graph (vertex and arc counts
and a CRC of the CSR arrays), the batch size and the window policy, so that a
later run loads the choice instead of searching again:

    opts = autotune.tuned_options(dg, g, q, cache="vine_amd/models/autotune.json",
                                  max_batch=4096, delta_policy=S.DELTA_MLP, ...)
    solver = S.Solver(dg, **opts)

Searched, one coordinate at a time from the library defaults: the engine
(one persistent workgroup per query, or grid-wide rounds) together with the
number of internal solvers on separate HIP streams (concurrency) and, for
grid-wide rounds, whether the rounds between host checks replay as one
captured HIP graph (use_graph); then, for grid-wide rounds only, frontier
entries a wave takes per expand step (labels_per_wave), threads per
workgroup of the round kernels (block_size), workgroups per round kernel
(grid_blocks) and the slots of the first-level candidate table in LDS
(hash_slots); last the Lagrangian multiplier mode of the bound searches
(lagrange_bound) and the near tier's reach in windows (near_windows).  Every configuration must return the same answers as the
first one (they only change the schedule); a difference raises.

    python -m vine_amd.autotune ny_batch        # tune the bench workload, update the cache
"""
import argparse
import json
import os
import sys
import time
import zlib

import numpy as np

from . import learn as L
from . import solver as S
from . import workload as W

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "autotune.json")
# coordinates searched one after another; a coordinate may set several options
# together (the engine, streams and graph replay interact: replay removes the
# launch gaps that other streams' kernels otherwise fill, and the workgroup
# engine has no launches per round to replay)
SPACE = ((("engine", "concurrency", "use_graph"),
          ((S.ENGINE_WORKGROUP, 1, 0), (S.ENGINE_WORKGROUP, 2, 0), (S.ENGINE_WORKGROUP, 3, 0),
           (S.ENGINE_ROUNDS, 1, 0), (S.ENGINE_ROUNDS, 1, 1), (S.ENGINE_ROUNDS, 2, 0), (S.ENGINE_ROUNDS, 2, 1),
           (S.ENGINE_ROUNDS, 3, 0), (S.ENGINE_ROUNDS, 3, 1))),
         (("labels_per_wave",), ((16,), (32,), (64,))),
         (("block_size",), ((64,), (128,), (256,))),
         (("grid_blocks",), ((256,), (384,), (768,))),
         (("hash_slots",), ((0,), (64,), (256,), (1024,))),
         (("lagrange_bound",), ((2,), (1,), (4,))),
         (("near_windows",), ((1.0,), (0.5,), (0.25,))))
# options of the grid-wide round kernels only (not searched for the workgroup engine)
ROUNDS_ONLY = ("labels_per_wave", "block_size", "grid_blocks", "hash_slots")
KEYS = tuple(k for names, _ in SPACE for k in names)
LIBRARY_DEFAULTS = {"engine": S.ENGINE_AUTO, "concurrency": 1, "labels_per_wave": 64, "block_size": 128,
                    "grid_blocks": 384, "use_graph": 0, "hash_slots": 0, "lagrange_bound": 2, "near_windows": 1.0}


def applies(names, opts):
    """Whether the coordinate `names` changes anything under `opts`."""
    return not (opts.get("engine") == S.ENGINE_WORKGROUP and all(k in ROUNDS_ONLY for k in names))


def graph_key(g, max_batch, policy):
    crc = 0
    for a in (g.row_ptr, g.dst, g.len, g.delay):
        crc = zlib.crc32(np.ascontiguousarray(a).tobytes(), crc)
    return "n%d_m%d_crc%08x_b%d_p%d" % (g.n, g.m, crc, max_batch, policy)


def _load(cache):
    if cache and os.path.exists(cache):
        with open(cache) as f:
            return json.load(f)
    return {}


def _time(dg, q, opts, mlp, reps):
    with S.Solver(dg, **opts) as s:
        if mlp is not None:
            s.set_mlp(mlp)
        a = s.solve(q.src, q.dst, q.bound, paths=False)  # warm-up
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            a = s.solve(q.src, q.dst, q.bound, paths=False)
            best = min(best, time.perf_counter() - t0)
    return best, a


def search(dg, q, base, mlp=None, reps=1, log=None):
    """Coordinate search over SPACE starting from `base` (dict of options).
    Returns (best options, queries/s, list of tried (options, queries/s))."""
    cur = dict(LIBRARY_DEFAULTS, **base)
    ref = None
    tried = []
    best_t = None
    for names, values in SPACE:
        if not applies(names, cur):
            continue
        scores = {}
        for v in values:
            opts = dict(cur, **dict(zip(names, v)))
            t, a = _time(dg, q, opts, mlp, reps)
            if ref is None:
                ref = a
            elif not (np.array_equal(ref.cost, a.cost) and np.array_equal(ref.delay, a.delay)
                      and np.array_equal(ref.feasible, a.feasible)):
                raise RuntimeError("launch configuration %r changed the answers" % (opts,))
            scores[v] = t
            tried.append(({k: opts[k] for k in KEYS}, len(q) / t))
            if log:
                log("%s=%s: %.1f queries/s" % ("/".join(names), "/".join(map(str, v)), len(q) / t))
        v_best = min(scores, key=scores.get)
        cur.update(zip(names, v_best))
        best_t = scores[v_best]
    return cur, len(q) / best_t, tried


def cached_options(g, cache=DEFAULT_CACHE, **base):
    """The cached launch options for this graph, batch size and policy, or None."""
    key = graph_key(g, base.get("max_batch", 1), base.get("delta_policy", S.DELTA_INF))
    entry = _load(cache).get(key)
    return dict(entry["options"]) if entry else None


def tuned_options(dg, g, q, cache=DEFAULT_CACHE, mlp=None, reps=1, retune=False, log=None, **base):
    """Options for Solver(dg, **options): the cached choice for this graph,
    batch size and policy, else a fresh search whose result is cached."""
    key = graph_key(g, base.get("max_batch", 1), base.get("delta_policy", S.DELTA_INF))
    entries = _load(cache)
    if key in entries and not retune:
        return dict(base, **entries[key]["options"])
    opts, qps, tried = search(dg, q, base, mlp=mlp, reps=reps, log=log)
    if cache:
        entries[key] = {"options": {k: opts[k] for k in KEYS}, "queries_per_s": qps, "queries": len(q),
                        "tried": [{"options": o, "queries_per_s": v} for o, v in tried]}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(entries, f, indent=1, sort_keys=True)
    return opts


BENCH_CAPACITIES = dict(label_capacity=1_000_000_000, pool_capacity=2_000_000_000, cand_capacity=100_000_000,
                        frontier_capacity=400_000_000)


def bench_base(batch=4096):
    """The batched, learned-window configuration bench.py runs."""
    return dict(max_batch=batch, prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_MLP, delta_fixed=4.0,
                rounds_per_sync=16, **BENCH_CAPACITIES)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["ny_batch"])
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    ap.add_argument("--reps", type=int, default=1)
    args = ap.parse_args(argv)
    g, q = W.preset(args.workload)[:2]
    dg = S.DeviceGraph(g)
    mlp = np.load(L.DEFAULT_MODEL)
    log = lambda m: print("[autotune] " + m, file=sys.stderr, flush=True)
    opts = tuned_options(dg, g, q, cache=args.cache, mlp=mlp, reps=args.reps, retune=True, log=log,
                         **bench_base(len(q)))
    print(json.dumps({k: opts[k] for k in KEYS}))


if __name__ == "__main__":
    main()
