"""Trains the learned join-order planner on the GPU.

Samples are cardinalities the matcher itself measures: seeded random-walk
queries (4-24 vertices; half-induced, induced and tree walks) on Chung-Lu
graphs of 100K, 300K and 1M vertices (gamma 2.1, average degree 16, 16
Zipf(1.0) labels), and cycles, cliques and induced walks of 4-8 vertices on
R-MAT graphs of 2^16 and 2^17 vertices (average degree 32, 32 Zipf(0.8)
labels), all with other seeds than the benchmark graphs, are matched with
their heuristic order, with every order the planner ranks (all connected
orders of a small query, else the greedy continuations) and with random
connected vertex orders; every join level of every completed run (within
the work budget) gives one (features, log1p(rows)) pair through
gpsm_plan_features.  The network is
trained by the device's own BF16-MFMA forward/backward kernels with Adam
(gpsm_mlp_train_step); the FP32 feature standardisation is saved beside the
weights and applied inside the kernels.

    python -m gpsm_amd.train_planner --samples 60000
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
from gpsm_amd.planner import MLP, WEIGHTS  # noqa: E402

OFF_W1, OFF_B1 = 0, 128 * 32


def random_order(q, rng):
    adj = [set() for _ in range(q.n)]
    for a, b in zip(q.src, q.dst):
        adj[int(a)].add(int(b))
        adj[int(b)].add(int(a))
    start = int(rng.integers(q.n))
    order, seen = [start], {start}
    while len(order) < q.n:
        front = sorted({w for v in order for w in adj[v]} - seen)
        w = int(front[rng.integers(len(front))])
        order.append(w)
        seen.add(w)
    return order


ALL_ORDERS = 64  # kAllOrders in plan.h


def connected_orders(n, ef, et, cap=ALL_ORDERS):
    """Every vertex order with connected prefixes, lexicographic; None if
    there are more than cap (plan.cpp all_connected_orders)."""
    adj = [set() for _ in range(n)]
    for a, b in zip(ef, et):
        adj[int(a)].add(int(b))
        adj[int(b)].add(int(a))
    out, od, placed = [], [], [False] * n

    def rec():
        if len(od) == n:
            if len(out) == cap:
                return False
            out.append(list(od))
            return True
        for w in range(n):
            if placed[w] or (od and not any(v in adj[w] for v in od)):
                continue
            placed[w] = True
            od.append(w)
            ok = rec()
            od.pop()
            placed[w] = False
            if not ok:
                return False
        return True

    return out if rec() else None


def candidate_list(q, ef, et, size):
    """The orders the device planner ranks besides the heuristic one
    (plan.cpp candidate_orders): all connected orders of a small query, else
    the greedy continuations."""
    allo = connected_orders(q.n, ef, et)
    return allo if allo is not None else continuations(q, ef, et, size)


def continuations(q, ef, et, size, first=4):
    """The candidate orders the device planner ranks (plan.cpp
    candidate_orders): from each of the `first` smallest-|CET| edges in both
    orientations, a min-|CET| and a most-constrained greedy continuation."""
    out = []
    for e in sorted(range(len(ef)), key=lambda i: (int(size[i]), i))[:first]:
        for a, b in ((int(ef[e]), int(et[e])), (int(et[e]), int(ef[e]))):
            for constrained in (False, True):
                od, placed = [a, b], {a, b}
                while len(od) < q.n:
                    best = None
                    for w in range(q.n):
                        if w in placed:
                            continue
                        hits = [int(size[i]) for i in range(len(ef))
                                if (ef[i] == w and et[i] in placed) or (et[i] == w and ef[i] in placed)]
                        if not hits:
                            continue
                        key = (-len(hits), min(hits)) if constrained else (min(hits),)
                        if best is None or key < best[0]:
                            best = (key, w)
                    od.append(best[1])
                    placed.add(best[1])
                if od not in out:
                    out.append(od)
    return out


# The training graphs, cycled over: Chung-Lu (gamma 2.1, average degree
# 16, 16 Zipf(1.0) labels) of three sizes with walk queries, and R-MAT
# (a, b, c = .57, .19, .19, average degree 32, 32 Zipf(0.8) labels) of two
# scales with cycles, cliques and induced walks; other seeds than the
# benchmark graphs.
SPECS = (("chunglu", 100_000), ("rmat", 16), ("chunglu", 300_000), ("rmat", 17), ("chunglu", 1_000_000))
KINDS = (_lib.Q_WALK_HALF, _lib.Q_WALK_HALF, _lib.Q_WALK_INDUCED, _lib.Q_WALK_TREE)
RMAT_KINDS = ((_lib.Q_CYCLE, 4, 6), (_lib.Q_CLIQUE, 4, 6), (_lib.Q_WALK_INDUCED, 4, 8))


def training_queries(gi, seed, queries_per_graph, budget, kmax):
    """(graph, queries) of training graph gi."""
    kind, size = SPECS[gi % len(SPECS)]
    if kind == "chunglu":
        g = G.chung_lu(size, 8 * size, 2.1, seed)
        G.set_labels(g, 16, 1.0, seed)
        return g, G.queries(g, queries_per_graph, 4, kmax, KINDS[(gi // len(SPECS)) % len(KINDS)], seed, budget)
    g = G.rmat(size, 16 << size, 0.57, 0.19, 0.19, seed)
    G.set_labels(g, 32, 0.8, seed)
    qs = []
    for j, (qk, lo, hi) in enumerate(RMAT_KINDS):
        qs += G.queries(g, queries_per_graph // len(RMAT_KINDS), lo, hi, qk, seed + 7 * j, 0)
    return g, qs


def collect(ctx, target_rows, seeds, queries_per_graph, budget, orders_per_query, work_budget, log, kmax=24):
    X, Y = [], []
    rng = np.random.default_rng(12345)
    for gi, seed in enumerate(seeds):
        g, qs = training_queries(gi, seed, queries_per_graph, budget, kmax)
        n = g.n
        kind = SPECS[gi % len(SPECS)][0]
        rg = ctx.upload(g)
        # queries whose heuristic join exceeds the work budget are left out
        base = rg.match_batch(qs, work_budget=work_budget)
        keep = [i for i, r in enumerate(base) if r.status == 0]
        qs = [qs[i] for i in keep]
        base = [base[i] for i in keep]
        tabs = [rg.cet_edges(q) for q in qs]
        # every (query, order) pair is one entry of a batch
        pairs = [(i, list(r.order)) for i, r in enumerate(base)]
        for i, q in enumerate(qs):
            if q.m == 0:
                continue
            pairs += [(i, od) for od in candidate_list(q, *tabs[i]) if od != list(base[i].order)]
            pairs += [(i, random_order(q, rng)) for _ in range(orders_per_query)]
        used = 0
        for c0 in range(0, len(pairs), 1024):
            chunk = pairs[c0:c0 + 1024]
            res = rg.match_batch([qs[i] for i, _ in chunk], explicit_orders=[od for _, od in chunk],
                                 work_budget=work_budget)
            for (i, od), r in zip(chunk, res):
                q = qs[i]
                if r.status != 0:
                    continue  # over the work budget
                assert (r.count, r.hash) == (base[i].count, base[i].hash)
                if r.count == 0 or q.n < 3:
                    continue
                f, t = rg.plan_features(q, r.cand, tabs[i][2], od, r.level_rows)
                X.append(f)
                Y.append(t)
                used += 1
        rg.free()
        rows = sum(len(y) for y in Y)
        log("graph seed %d (%s, n=%d): %d queries, %d runs, %d rows so far"
            % (seed, kind, n, len(qs), used, rows))
        if rows >= target_rows:
            break
    return np.concatenate(X), np.concatenate(Y)


def train(ctx, X, Y, epochs, lr, log, seed=0):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    nval = max(len(X) // 10, 1)
    val, tr = perm[:nval], perm[nval:]
    mu = X[tr].mean(0).astype(np.float32)
    sd = X[tr].std(0)
    sd[sd < 1e-6] = 1.0
    isd = (1.0 / sd).astype(np.float32)
    Xn = X.astype(np.float32)  # standardised inside the kernels (FP32, before BF16 rounding)
    m = MLP(ctx, seed=seed)
    m.set_norm(mu, isd)
    hist = []
    for ep in range(epochs):
        idx = rng.permutation(tr)
        cur_lr = lr * (0.5 ** (ep // max(epochs // 4, 1)))
        for i in range(0, len(idx), 256):
            b = idx[i:i + 256]
            m.train_step(Xn[b], Y[b], cur_lr)
        if ep % 5 == 4 or ep == epochs - 1:
            pv = m.forward(Xn[val])
            pt = m.forward(Xn[tr[:4096]])
            hist.append({"epoch": ep + 1, "val_mse": float(np.mean((pv - Y[val]) ** 2)),
                         "train_mse": float(np.mean((pt - Y[tr[:4096]]) ** 2))})
            log("epoch %d: train mse %.4f  val mse %.4f" % (ep + 1, hist[-1]["train_mse"], hist[-1]["val_mse"]))
    p = m.get().copy()
    baseline = float(np.mean((Y[val] - Y[tr].mean()) ** 2))
    m.free()
    return p, np.concatenate([mu, isd]).astype(np.float32), hist, baseline


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=60000)
    ap.add_argument("--queries-per-graph", type=int, default=300)
    ap.add_argument("--orders", type=int, default=3)
    ap.add_argument("--budget", type=float, default=3e7)
    ap.add_argument("--work-budget", type=float, default=4e8)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    t0 = time.time()

    def log(s):
        print("[%6.1fs] %s" % (time.time() - t0, s), flush=True)

    ctx = Context(0)
    X, Y = collect(ctx, args.samples, range(101, 160), args.queries_per_graph, int(args.budget), args.orders,
                   int(args.work_budget), log)
    log("collected %d samples" % len(X))
    p, norm, hist, baseline = train(ctx, X, Y, args.epochs, args.lr, log)
    # the saved state reproduces the trained network on raw features
    m = MLP(ctx)
    m.set(p)
    m.set_norm(norm[:32], norm[32:])
    pred = m.forward(X[:4096])
    log("saved network mse on raw features: %.4f (constant predictor %.4f)"
        % (float(np.mean((pred - Y[:4096]) ** 2)), baseline))
    np.save(args.out, p.astype(np.float32), allow_pickle=False)
    np.save(os.path.join(os.path.dirname(args.out), "planner_norm.npy"), norm, allow_pickle=False)
    meta = {"samples": int(len(X)), "history": hist, "constant_predictor_val_mse": baseline,
            "device": ctx.device_name, "args": vars(args), "seconds": round(time.time() - t0, 1)}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log("saved %s" % args.out)


if __name__ == "__main__":
    main()
