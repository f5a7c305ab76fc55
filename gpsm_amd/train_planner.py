"""Trains the learned join-order planner on the GPU.

Samples are cardinalities the matcher itself measures: seeded random-walk
queries on Chung-Lu graphs with config-2 parameters (gamma 2.1, 16 Zipf(1.0)
labels, other seeds than the benchmark graph) are matched with their
heuristic order and with random connected vertex orders; every join level
of every run gives one (features, log1p(rows)) pair through
gpsm_plan_features.  The network is trained by the device's own BF16-MFMA
forward/backward kernels with Adam (gpsm_mlp_train_step), and the feature
standardisation is folded into the first layer before saving.

    python -m gpsm_amd.train_planner --samples 20000
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


def collect(ctx, target_rows, seeds, n, m, queries_per_graph, budget, orders_per_query, log):
    X, Y = [], []
    rng = np.random.default_rng(12345)
    for seed in seeds:
        g = G.chung_lu(n, m, 2.1, seed)
        G.set_labels(g, 16, 1.0, seed)
        qs = G.queries(g, queries_per_graph, 4, 10, _lib.Q_WALK_HALF, seed, budget)
        rg = ctx.upload(g)
        base = rg.match_batch(qs)
        cets = [rg.cet_sizes(q) for q in qs]
        runs = [(list(r.order), r) for r in base]
        for rep in range(orders_per_query):
            ords = [random_order(q, rng) for q in qs]
            res = rg.match_batch(qs, explicit_orders=ords)
            for q, r0, r in zip(qs, base, res):
                assert (r.count, r.hash) == (r0.count, r0.hash)
            runs += list(zip(ords, res))
        for i, (od, r) in enumerate(runs):
            q = qs[i % len(qs)]
            if r.count == 0 or q.n < 3:
                continue
            f, t = rg.plan_features(q, r.cand, cets[i % len(qs)], od, r.level_rows)
            X.append(f)
            Y.append(t)
        rg.free()
        rows = sum(len(y) for y in Y)
        log("graph seed %d: %d queries, %d rows so far" % (seed, len(qs), rows))
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
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--m", type=int, default=8_000_000)
    ap.add_argument("--queries-per-graph", type=int, default=400)
    ap.add_argument("--orders", type=int, default=5)
    ap.add_argument("--budget", type=float, default=3e7)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    t0 = time.time()

    def log(s):
        print("[%6.1fs] %s" % (time.time() - t0, s), flush=True)

    ctx = Context(0)
    X, Y = collect(ctx, args.samples, range(101, 140), args.n, args.m, args.queries_per_graph, int(args.budget),
                   args.orders, log)
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
