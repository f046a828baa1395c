"""Run a named workload at full size under several window policies.

Every policy must give the same answers (windows only reorder work), which
is checked here; the oracle comparison of these workloads lives in tests/.
Prints one JSON line per policy: queries/s, rounds, labels expanded, p50/p99
latency, and whether the run overflowed the configured capacities.

    python -m vine_amd.evaluate ne --policies mlp fixed:12 ar --queries 1000
    python -m vine_amd.evaluate rmat17 --policies fixed:8 inf

Workloads (vine_amd.workload.preset): grid64 (C1), ny (C2, bound 1.1x),
ne (C3: 1,524,453-vertex 3-NN graph, 1,000 random pairs, bound 1.2x),
rmat17 (C4), ny_batch (C5).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import learn as L
from . import solver as S
from . import workload as W


def parse_policy(text):
    """(policy, window, AR MAX or None); ar:<MAX> and dop:<MAX> set the rule's MAX."""
    if text == "mlp":
        return S.DELTA_MLP, 4.0, None
    for name, pol in (("ar", S.DELTA_AR), ("dop", S.DELTA_DOP)):
        if text == name:
            return pol, 0.0, None
        if text.startswith(name + ":"):
            return pol, 0.0, float(text.split(":", 1)[1])
    if text == "inf":
        return S.DELTA_INF, 0.0, None
    if text.startswith("fixed:"):
        return S.DELTA_FIXED, float(text.split(":", 1)[1]), None
    raise ValueError("policy must be mlp, ar[:MAX], dop[:MAX], inf or fixed:<window>")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload")
    ap.add_argument("--policies", nargs="+", default=["mlp", "fixed:12"])
    ap.add_argument("--queries", type=int, default=0, help="first N queries (0 = all)")
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--model", default=L.DEFAULT_MODEL)
    ap.add_argument("--engine", type=int, default=S.ENGINE_AUTO)
    ap.add_argument("--concurrency", type=int, default=1, help="internal solvers on separate HIP streams")
    ap.add_argument("--label-capacity", type=float, default=6e8)
    ap.add_argument("--ar-max-labels", type=float, default=0.0,
                    help="MAX of the AR rule (0 = the reference's: free memory / 12 B)")
    ap.add_argument("--chunk", type=int, default=0, help="solve the queries in calls of this many (progress lines)")
    args = ap.parse_args(argv)

    t0 = time.time()
    got = W.preset(args.workload)
    g, qs = got[0], got[1]
    if args.queries:
        qs = qs.subset(np.arange(min(args.queries, len(qs))))
    print("[evaluate] %s: n=%d m=%d, %d queries (%.1f s to generate)" % (args.workload, g.n, g.m, len(qs),
                                                                       time.time() - t0), file=sys.stderr, flush=True)
    dg = S.DeviceGraph(g)
    params = np.load(args.model) if os.path.exists(args.model) else None
    ref = None
    for text in args.policies:
        policy, delta, mx = parse_policy(text)
        cap = int(args.label_capacity)
        line = {"workload": args.workload, "policy": text, "queries": len(qs), "batch": args.batch,
                "concurrency": args.concurrency}
        try:
            with S.Solver(dg, max_batch=args.batch, prune=S.PRUNE_BOUNDS, delta_policy=policy, delta_fixed=delta,
                          rounds_per_sync=16 if policy not in (S.DELTA_AR, S.DELTA_DOP) else 1,
                          engine=args.engine,
                          concurrency=args.concurrency,
                          label_capacity=cap, pool_capacity=2 * cap, cand_capacity=100_000_000,
                          frontier_capacity=cap // 2,
                          ar_max_labels=args.ar_max_labels if mx is None else mx) as s:
                if policy == S.DELTA_MLP:
                    if params is None:
                        raise RuntimeError("no model at %s" % args.model)
                    s.set_mlp(params)
                t1 = time.perf_counter()
                if args.chunk:
                    parts, lat = [], []
                    for c0 in range(0, len(qs), args.chunk):
                        sub = qs.subset(np.arange(c0, min(c0 + args.chunk, len(qs))))
                        parts.append(s.solve(sub.src, sub.dst, sub.bound, paths=False))
                        lat.append(s.latency_us(len(sub)))
                        print("[evaluate] %s: %d/%d queries, %.1f s" % (text, min(c0 + args.chunk, len(qs)), len(qs),
                                                                       time.perf_counter() - t1),
                              file=sys.stderr, flush=True)
                    a = S.Answer(np.concatenate([p.cost for p in parts]), np.concatenate([p.delay for p in parts]),
                                 np.concatenate([p.feasible for p in parts]), None, None,
                                 {k: sum(p.stats[k] for p in parts) for k in parts[0].stats})
                    lat = np.concatenate(lat)
                else:
                    a = s.solve(qs.src, qs.dst, qs.bound, paths=False)
                    lat = s.latency_us(len(qs))
                dt = time.perf_counter() - t1
            st = a.stats
            line.update({"seconds": dt, "qps": len(qs) / dt, "rounds": st["rounds"], "expanded": st["expanded"],
                         "inserted": st["inserted"], "feasible": int(a.feasible.sum()),
                         "p50_ms": float(np.percentile(lat, 50)) / 1e3, "p99_ms": float(np.percentile(lat, 99)) / 1e3,
                         "status": "ok"})
            if ref is None:
                ref = (a.cost.copy(), a.delay.copy(), a.feasible.copy(), text)
            else:
                same = (np.array_equal(ref[0], a.cost) and np.array_equal(ref[1], a.delay)
                        and np.array_equal(ref[2], a.feasible))
                line["same_answers_as"] = ref[3]
                line["same_answers"] = bool(same)
        except S.VineError as e:
            line.update({"status": "error", "error": str(e)})
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
