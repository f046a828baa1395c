"""Training the bucket-width MLP from traces the solver generates itself.

1. Training queries (seeds disjoint from the evaluation seeds) are solved by
   the batched solver once per candidate fixed window; per query the solver
   reports labels expanded and active rounds, and per (query, round) the eight
   scheduler features.
2. Each query is labelled with a window (label_windows): rounds are shared by
   the batch, so a round budget is chosen to minimise labels expanded +
   round_cost * rounds over the batch, and each query takes its least-work
   window within the budget.  ROUND_COST (how many label expansions one
   round's fixed kernel time is worth) was chosen on queries of a separate
   validation seed.  With --wide the heaviest queries (the ones that set the
   batch's round count) are also run under wide windows, so the labels can
   give them wide windows (measured: no gain on the NY workload, the network
   widens too many queries).  Every feature row recorded under that window
   becomes a sample with target
   log2(window); the second output is trained to forecast log2(1 +
   expansions) TRACE_EVERY rounds ahead from the same rows.
3. The network is trained on the GPU (vine_mlp_train: HIP forward, backward
   and Adam kernels) and saved as an .npy parameter vector.

Answers never depend on the network: windows only reorder work.

    python -m vine_amd.learn --out vine_amd/models/delta_mlp.npy
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import mlp as M
from . import solver as S
from . import workload as W

CANDIDATES = (4.0, 8.0, 12.0, 16.0, 24.0, 32.0, 48.0)
WIDE = (64.0, 96.0, 128.0, 192.0, 256.0)  # run on the heaviest queries only
HEAVY_FRACTION = 0.01  # queries with the most rounds under the widest base window
# labels one round's fixed kernel time is worth (chosen among 400, 800,
# 1,150 and 1,800 on a validation query seed)
ROUND_COST = 1150.0
TRACE_EVERY = 4  # feature rows every 4th round
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "delta_mlp.npy")


def log(msg):
    print("[learn] " + msg, file=sys.stderr, flush=True)


LAGRANGE = None  # multiplier searches of the bound searches (None: the library default)


def solver_for(dg, batch, policy, delta, trace=0, **kw):
    extra = {} if LAGRANGE is None else {"lagrange_bound": LAGRANGE}
    return S.Solver(dg, max_batch=batch, prune=S.PRUNE_BOUNDS, delta_policy=policy, delta_fixed=delta,
                    rounds_per_sync=16, record_trace=trace, label_capacity=kw.get("label_capacity", 600_000_000),
                    pool_capacity=kw.get("pool_capacity", 1_200_000_000), cand_capacity=100_000_000,
                    frontier_capacity=400_000_000, **extra)


def collect(dg, q, batch, candidates=CANDIDATES):
    """Per candidate window: per-query (expanded, rounds) and the feature rows."""
    out = {}
    for d in candidates:
        t0 = time.time()
        with solver_for(dg, batch, S.DELTA_FIXED, d, trace=TRACE_EVERY) as s:
            a = s.solve(q.src, q.dst, q.bound, paths=False)
            ex, rd = s.query_stats(len(q))
            tr = s.trace()
        out[d] = {"expanded": ex, "rounds": rd, "trace": tr, "cost": a.cost, "feasible": a.feasible,
                  "seconds": time.time() - t0, "batch_rounds": int(a.stats["rounds"]),
                  "batch_expanded": int(a.stats["expanded"])}
        log("window %g: %.1f s, %d trace rows" % (d, time.time() - t0, tr.shape[0]))
    return out


def collect_wide(dg, q, runs, windows=WIDE, fraction=HEAVY_FRACTION):
    """The heaviest queries (most rounds under the widest base window) again
    under wide windows; other queries get +inf work and rounds there, so the
    labels never pick a wide window for them."""
    base = runs[max(runs)]["rounds"]
    k = max(1, int(len(q) * fraction))
    heavy = np.sort(np.argsort(-base)[:k])
    sub = q.subset(heavy)
    log("heavy queries: %d, rounds under window %g: %d .. %d" % (k, max(runs), base[heavy].min(), base[heavy].max()))
    for d in windows:
        t0 = time.time()
        ex = np.full(len(q), np.inf)
        rd = np.full(len(q), np.inf)
        try:
            with solver_for(dg, len(sub), S.DELTA_FIXED, d, trace=TRACE_EVERY,
                            label_capacity=1_500_000_000, pool_capacity=3_000_000_000) as s:
                s.solve(sub.src, sub.dst, sub.bound, paths=False)
                e2, r2 = s.query_stats(len(sub))
                tr = s.trace()
        except S.VineError as e:
            log("window %g on the heavy queries: %s; skipped" % (d, e))
            continue
        ex[heavy] = e2
        rd[heavy] = r2
        if tr.shape[0]:
            tr = tr.copy()
            tr[:, 0] = heavy[tr[:, 0].astype(np.int64)]
        runs[d] = {"expanded": ex, "rounds": rd, "trace": tr, "seconds": time.time() - t0}
        log("window %g (heavy): %.1f s, rounds max %d, expanded %.3g" % (d, time.time() - t0, r2.max(), e2.sum()))
    return runs


def label_windows(E, R, round_cost=ROUND_COST):
    """Batch-aware window labels.  Rounds are shared by the whole batch, so
    only the queries that set the round count pay for their rounds: for a
    round budget Rb every query takes its least-work window finishing within
    Rb (or its fewest-rounds window), and Rb minimises
    total labels expanded + round_cost * rounds of the batch.
    E, R: [queries, windows].  Returns (labels, rounds of the batch, cost)."""
    budgets = np.unique(np.quantile(R.min(1), np.linspace(0.5, 1.0, 41)).astype(np.int64))
    budgets = np.unique(np.concatenate([budgets, [R.min(1).max(), R.max()]]))
    best = None
    for rb in budgets:
        ok = R <= rb
        pick = np.where(ok.any(1), np.argmin(np.where(ok, E, np.inf), 1), np.argmin(R, 1))
        rows = np.arange(E.shape[0])
        rounds = R[rows, pick].max()
        total = E[rows, pick].sum() + round_cost * rounds
        if best is None or total < best[2]:
            best = (pick, int(rounds), float(total))
    return best


def rows_to_samples(tr, window_of_query, nq):
    """Feature rows -> (X, T, mask): target 0 is log2 of the query's window,
    target 1 the expansions at the query's next recorded round."""
    qi = tr[:, 0].astype(np.int64)
    rnd = tr[:, 1].astype(np.int64)
    keep = qi < nq
    qi, rnd, tr = qi[keep], rnd[keep], tr[keep]
    order = np.lexsort((rnd, qi))
    qs, xs = qi[order], tr[order, 2:]
    nxt = np.full(qs.shape[0], np.nan, np.float32)
    same = qs[1:] == qs[:-1]
    nxt[:-1][same] = xs[1:, 2][same]
    t = np.zeros((qs.shape[0], 2), np.float32)
    t[:, 0] = np.log2(window_of_query[qs])
    t[:, 1] = np.nan_to_num(nxt)
    m = np.ones_like(t)
    m[:, 1] = np.isfinite(nxt).astype(np.float32)
    return xs, t, m


def build_dataset(runs, batch, nq, max_rows=600_000, seed=0, round_cost=ROUND_COST):
    ds = sorted(runs)
    E = np.stack([runs[d]["expanded"] for d in ds], 1).astype(np.float64)
    R = np.stack([runs[d]["rounds"] for d in ds], 1).astype(np.float64)
    best, budget, cost = label_windows(E, R, round_cost)
    log("round budget %d, modelled cost %.3g" % (budget, cost))
    X, T, Mk = [], [], []
    window_of_query = np.asarray(ds, np.float32)[best]
    for k, d in enumerate(ds):
        tr = runs[d]["trace"]
        if tr.shape[0] == 0:
            continue
        qi = tr[:, 0].astype(np.int64)
        sel = (qi < nq) & (best[np.minimum(qi, nq - 1)] == k)
        x, t, m = rows_to_samples(tr[sel], window_of_query, nq)
        X.append(x)
        T.append(t)
        Mk.append(m)
    X = np.concatenate(X)
    T = np.concatenate(T)
    Mk = np.concatenate(Mk)
    if X.shape[0] > max_rows:
        idx = np.random.default_rng(seed).choice(X.shape[0], max_rows, replace=False)
        X, T, Mk = X[idx], T[idx], Mk[idx]
    return X, T, Mk, best, cost


def train(X, T, Mk, epochs=20, lr=2e-3, seed=0):
    sd = X.std(0)
    sd[sd < 1e-3] = 1.0
    net = M.MLP(M.init_params(seed, xmu=X.mean(0), xsd=sd))
    curve = net.train(X, T, Mk, epochs=epochs, batch=256, lr=lr, seed=seed)
    return net.params(), curve


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--queries", type=int, default=8192)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1000, help="query seed (evaluation uses 0)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--dagger", type=int, default=2)
    ap.add_argument("--wide", action="store_true", help="also offer wide windows to the heaviest queries")
    ap.add_argument("--lagrange-bound", type=int, default=None,
                    help="multiplier searches of the bound searches (default: the library's)")
    args = ap.parse_args()
    global LAGRANGE
    LAGRANGE = args.lagrange_bound
    g = W.knn(264346, seed=0)
    q = W.hop_queries(g, args.queries, seed=args.seed)
    log("training queries: %d (seed %d)" % (len(q), args.seed))
    dg = S.DeviceGraph(g)
    runs = collect(dg, q, args.batch)
    if args.wide:
        runs = collect_wide(dg, q, runs)
    round_cost = ROUND_COST
    X, T, Mk, best, cost = build_dataset(runs, args.batch, len(q), round_cost=round_cost)
    wins = sorted(runs)
    log("dataset: %d rows; best-window histogram %s over %s" % (X.shape[0], np.bincount(best, minlength=len(wins)).tolist(), wins))
    params, curve = train(X, T, Mk, epochs=args.epochs)
    # DAgger-style passes: the network's own trajectories, labelled with each
    # query's window, so it also learns the states it reaches itself
    window_of_query = np.asarray(wins, np.float32)[best]
    for it in range(args.dagger):
        t0 = time.time()
        try:
            with solver_for(dg, args.batch, S.DELTA_MLP, CANDIDATES[0], trace=TRACE_EVERY) as s:
                s.set_mlp(params)
                s.solve(q.src, q.dst, q.bound, paths=False)
                tr = s.trace()
        except S.VineError as e:
            log("dagger %d: on-policy run failed (%s); stopping" % (it, e))
            break
        x2, t2, m2 = rows_to_samples(tr, window_of_query, len(q))
        X, T, Mk = np.concatenate([X, x2]), np.concatenate([T, t2]), np.concatenate([Mk, m2])
        log("dagger %d: %.1f s on-policy, %d rows added" % (it, time.time() - t0, x2.shape[0]))
        params, curve = train(X, T, Mk, epochs=args.epochs, seed=it + 1)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, params)
    meta = {"queries": len(q), "seed": args.seed, "candidates": wins, "round_cost": round_cost,
            "lagrange_bound": LAGRANGE,
            "rows": int(X.shape[0]), "loss_curve": [float(c) for c in curve],
            "best_hist": [int(x) for x in np.bincount(best, minlength=len(wins))],
            "window_seconds": {str(d): runs[d]["seconds"] for d in runs}}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log("saved %s; loss %.4f -> %.4f" % (args.out, curve[0], curve[-1]))


if __name__ == "__main__":
    main()
