"""Measures what the join's time is made of, for the planner's cost model.

The planner ranks candidate join orders by a predicted cost in units of one
examined candidate.  Examined candidates are not all alike: a candidate of a
level with b join-mode (back) edges costs b edge-membership searches more,
and every materialised row costs a window load (its columns, a rank-index
lookup and an offset pair).  This script measures the weights on the device.

Batches of cycles and cliques of 4-6 vertices (every vertex after the first
edge has back edges) on an R-MAT graph of 2^18 vertices (a, b, c = .57,
.19, .19, average degree 32, 32 Zipf(0.8) labels), kept by the device-exact
feasibility rule at a work budget of 5e8, are matched with explicit orders,
every query taking one of the orders the planner ranks (the heuristic order
and candidate_list's) at random.  Per batch the join kernel's time (HIP
events, best of 3) and, from the device's per-level counters,

    W1  first-table rows
    W   candidates examined at levels >= 2
    WB  candidates examined times the back edges of their level
    R   rows materialised at levels 2 .. q-2

are recorded, and the time is fitted as a non-negative combination of them.
The weights relative to W are the cost model's constants (plan.h).

    python -m gpsm_amd.calibrate_planner --batches 40 --out calib.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from gpsm_amd import _lib  # noqa: E402
from gpsm_amd import graphs as G  # noqa: E402
from gpsm_amd.matcher import Context  # noqa: E402
from gpsm_amd.train_planner import candidate_list  # noqa: E402
from gpsm_amd.workloads import feasible_queries  # noqa: E402


def backs_per_level(q, order, ef, et, size):
    """Join-mode edges of every position (edges to earlier vertices other
    than the expand edge), as the device's plan has them."""
    pos = {int(v): k for k, v in enumerate(order)}
    ef, et = [int(x) for x in ef], [int(x) for x in et]
    out = [0] * q.n
    for k in range(2, q.n):
        w = order[k]
        mine = [i for i in range(len(ef)) if (ef[i] == w and pos[et[i]] < k) or (et[i] == w and pos[ef[i]] < k)]
        others = {pos[int(et[i]) if ef[i] == w else int(ef[i])] for i in mine}
        out[k] = len(others) - 1
    return out


def features(q, r, backs):
    lw, lr = r.level_work, r.level_rows
    W1 = lw[1]
    W = sum(lw[2:q.n])
    WB = sum(lw[k] * backs[k] for k in range(2, q.n))
    R = sum(lr[2:q.n - 1])
    return [W1, W, WB, R]


def nnls(A, y):
    """Non-negative least squares by coordinate descent (small problems)."""
    x = np.zeros(A.shape[1])
    AtA, Aty = A.T @ A, A.T @ y
    for _ in range(20000):
        for j in range(len(x)):
            if AtA[j, j] <= 0:
                continue
            x[j] = max(0.0, x[j] + (Aty[j] - AtA[j] @ x) / AtA[j, j])
    return x


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batches", type=int, default=24)
    ap.add_argument("--queries", type=int, default=32)
    ap.add_argument("--budget", type=float, default=5e8, help="work budget of the feasibility rule")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    t0 = time.time()

    def log(s):
        print("[%6.1fs] %s" % (time.time() - t0, s), flush=True)

    g = G.rmat(18, 16 << 18, 0.57, 0.19, 0.19, 31)
    G.set_labels(g, 32, 0.8, 31)
    ctx = Context(0)
    rg = ctx.upload(g)
    qs = []
    for kind, seed in ((_lib.Q_CYCLE, 41), (_lib.Q_CLIQUE, 42)):
        q, _, _ = feasible_queries(rg, args.queries // 2, 4, 6, [kind], seed, int(args.budget), per_round=6, rounds=3)
        qs += q
    base = rg.match_batch(qs)
    tabs = [rg.cet_edges(q) for q in qs]
    cands = []
    for q, r, (ef, et, size) in zip(qs, base, tabs):
        c = [list(r.order)]
        c += [od for od in candidate_list(q, ef, et, size) if od not in c]
        cands.append(c)
    log("%d queries, %.1f candidate orders each" % (len(qs), np.mean([len(c) for c in cands])))
    rng = np.random.default_rng(args.seed)
    X, y, picks = [], [], []
    for b in range(args.batches + 1):
        pick = [0] * len(qs) if b == 0 else [int(rng.integers(len(c))) for c in cands]
        orders = [cands[i][p] for i, p in enumerate(pick)]
        best = None
        for _ in range(3):
            ctx.profile(1)
            res = rg.match_batch(qs, explicit_orders=orders)
            ms = ctx.kernel_stats()["join"]["ms"]
            best = ms if best is None else min(best, ms)
        ctx.profile(0)
        assert [(r.count, r.hash) for r in res] == [(r.count, r.hash) for r in base]
        f = np.sum([features(q, r, backs_per_level(q, od, *tabs[i][:2], tabs[i][2]))
                    for i, (q, r, od) in enumerate(zip(qs, res, orders))], axis=0)
        X.append(f.tolist())
        y.append(best)
        picks.append(pick)
        log("batch %d: join %.2f ms, W1 %d W %d WB %d R %d" % (b, best, *f))
    A = np.array(X, dtype=np.float64)
    Y = np.array(y)
    scale = A.max(0)
    scale[scale == 0] = 1
    coef = nnls(A / scale, Y) / scale
    pred = A @ coef
    rel = {k: (coef[i] / coef[1] if coef[1] > 0 else None) for i, k in enumerate(["W1", "W", "WB", "R"])}
    log("ms per unit: W1 %.3g  W %.3g  WB %.3g  R %.3g" % tuple(coef))
    log("relative to an examined candidate: %s" % json.dumps(rel))
    log("fit: max |err| %.2f ms of %.2f-%.2f ms" % (np.abs(pred - Y).max(), Y.min(), Y.max()))
    out = {"ms_per_unit": dict(zip(["W1", "W", "WB", "R"], coef.tolist())), "relative": rel,
           "X": X, "join_ms": y, "device": ctx.device_name, "queries": len(qs), "batches": args.batches,
           "seconds": round(time.time() - t0, 1)}
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(out, fh, indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
