"""Time the label-family search (vine_family_solve) on a seeded k-NN graph.

    python -m vine_amd.evaluate_families --families csp hop qos --windows 40 400

Graph: 20,000 uniform points, 3-NN union-symmetrised (len = round(dist *
1e5), delay = round(len * U[0.5, 2])); 256 random pairs (seed 5), delay
bound floor(1.2 x min delay).  Limits per family: the delay bound; hop
families allow 4,000 arcs (and a longest arc of 10^6); QoS asks for a
bottleneck of at least the 20th percentile of the arc delays.  Prints one
JSON line per (family, window): seconds, queries/s, rounds, labels
expanded, feasible count, and for the families whose answer is the c_label
answer (csp, hop, hopmax with loose hop limits) whether every cost equals
the main solver's.
"""
import argparse
import json
import time

import numpy as np

from . import solver as S
from . import workload as W

FAMILIES = {"csp": S.FAMILY_CSP, "hop": S.FAMILY_HOP, "hopmax": S.FAMILY_HOP_MAXCOST, "2osp": S.FAMILY_2OSP,
            "qos": S.FAMILY_QOS}
SAME_AS_MAIN = ("csp", "hop", "hopmax", "2osp")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--families", nargs="+", default=["csp", "hop", "hopmax", "qos"], choices=sorted(FAMILIES))
    ap.add_argument("--windows", nargs="+", type=float, default=[40.0, 400.0])
    ap.add_argument("--vertices", type=int, default=20000)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--label-capacity", type=int, default=400_000_000)
    args = ap.parse_args(argv)

    g = W.knn(args.vertices, seed=0)
    q = W.random_queries(g, args.queries, (12, 10), seed=5)
    dg = S.DeviceGraph(g)
    with S.Solver(dg, max_batch=len(q), prune=S.PRUNE_BOUNDS, delta_policy=S.DELTA_FIXED, delta_fixed=40.0,
                  rounds_per_sync=16) as s:
        main_ans = s.solve(q.src, q.dst, q.bound, paths=False)
    for name in args.families:
        fam = FAMILIES[name]
        lim = np.zeros((len(q), 3), np.float32)
        lim[:, 0] = q.bound
        lim[:, 1] = 4000
        lim[:, 2] = 1e6
        if fam == S.FAMILY_QOS:
            lim[:, 0] = np.percentile(g.delay, 20)
        for window in args.windows:
            with S.FamilySolver(dg, fam, max_batch=args.batch, label_capacity=args.label_capacity, prune=1,
                                window=window) as fs:
                t0 = time.perf_counter()
                a = fs.solve(q.src, q.dst, lim[:, 0], lim[:, 1], lim[:, 2], paths=False)
                dt = time.perf_counter() - t0
            line = {"family": name, "window": window, "queries": len(q), "seconds": round(dt, 4),
                    "queries_per_s": round(len(q) / dt, 1), "rounds": a.counters["rounds"],
                    "expanded": a.counters["expanded"], "feasible": int(a.feasible.sum())}
            if name in SAME_AS_MAIN:
                line["same_cost_as_main"] = bool(np.array_equal(a.feasible, main_ans.feasible) and
                                                 np.array_equal(a.label[a.feasible, 0],
                                                                main_ans.cost[main_ans.feasible]))
            print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
