"""Training the bucket-width MLP from traces the solver generates itself.

1. Training queries (seeds disjoint from the evaluation seeds) are solved by
   the batched solver once per candidate fixed window; per query the solver
   reports labels expanded and active rounds, and per (query, round) the eight
   scheduler features.
2. Each query is labelled with a window (label_windows): rounds are shared by
   the batch, so a round budget is chosen to minimise labels expanded +
   round_cost * rounds over the batch, and each query takes its least-work
   window within the budget.  round_cost (how many label expansions one
   round's fixed kernel time is worth) is fitted by least squares to the
   collection runs themselves: seconds ~ a * rounds + b * expanded + c.  Every feature row recorded under that window
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

CANDIDATES = (4.0, 8.0, 12.0, 16.0, 24.0, 32.0, 48.0, 64.0, 96.0)
# fallback for round_cost when the fit is degenerate (labels a round's fixed
# kernel time is worth on MI355X: ~63 us per round, ~19 ns per expansion)
ROUND_COST = 3400.0
TRACE_EVERY = 4  # feature rows every 4th round
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "delta_mlp.npy")


def log(msg):
    print("[learn] " + msg, file=sys.stderr, flush=True)


def solver_for(dg, batch, policy, delta, trace=0, **kw):
    return S.Solver(dg, max_batch=batch, prune=S.PRUNE_BOUNDS, delta_policy=policy, delta_fixed=delta,
                    rounds_per_sync=16, record_trace=trace, label_capacity=kw.get("label_capacity", 600_000_000),
                    pool_capacity=kw.get("pool_capacity", 1_200_000_000), cand_capacity=100_000_000,
                    frontier_capacity=400_000_000)


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


def fit_round_cost(runs):
    """Least squares seconds ~ a * rounds + b * expanded + c over the
    collection runs; returns a / b (ROUND_COST if the fit is degenerate)."""
    ds = sorted(runs)
    if len(ds) < 3:
        return ROUND_COST
    A = np.array([[runs[d]["batch_rounds"], runs[d]["batch_expanded"], 1.0] for d in ds], np.float64)
    y = np.array([runs[d]["seconds"] for d in ds], np.float64)
    coef = np.linalg.lstsq(A, y, rcond=None)[0]
    if not (coef[0] > 0 and coef[1] > 0):
        return ROUND_COST
    return float(coef[0] / coef[1])


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
    args = ap.parse_args()
    g = W.knn(264346, seed=0)
    q = W.hop_queries(g, args.queries, seed=args.seed)
    log("training queries: %d (seed %d)" % (len(q), args.seed))
    dg = S.DeviceGraph(g)
    runs = collect(dg, q, args.batch)
    round_cost = fit_round_cost(runs)
    log("fitted round cost: %.0f label expansions per round" % round_cost)
    X, T, Mk, best, cost = build_dataset(runs, args.batch, len(q), round_cost=round_cost)
    log("dataset: %d rows; best-window histogram %s" % (X.shape[0], np.bincount(best, minlength=len(CANDIDATES))))
    params, curve = train(X, T, Mk, epochs=args.epochs)
    # DAgger-style passes: the network's own trajectories, labelled with each
    # query's window, so it also learns the states it reaches itself
    window_of_query = np.asarray(CANDIDATES, np.float32)[best]
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
    meta = {"queries": len(q), "seed": args.seed, "candidates": list(CANDIDATES), "round_cost": round_cost,
            "rows": int(X.shape[0]), "loss_curve": [float(c) for c in curve],
            "best_hist": [int(x) for x in np.bincount(best, minlength=len(CANDIDATES))],
            "window_seconds": {str(d): runs[d]["seconds"] for d in runs}}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log("saved %s; loss %.4f -> %.4f" % (args.out, curve[0], curve[-1]))


if __name__ == "__main__":
    main()
