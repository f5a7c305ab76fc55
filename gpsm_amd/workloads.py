"""The benchmark workloads, built from seeded generators.

* er10k_walk6: ER 10k vertices / 40k edges, 8 uniform labels, one 6-vertex
  walk query (the size the CPU enumerator finishes in milliseconds).
* chunglu1m_walk8_16: Chung-Lu 1M/8M (gamma 2.1), 16 Zipf(1.0) labels, 100
  walk queries drawn with 8-16 vertices, half tree-like, half induced.
* chunglu1m_mix4_32: the same graph with 256 walk queries drawn with 4-32
  vertices (matched with the heuristic or the learned join order).
* rmat22_64m + rmat_shape_queries_device: R-MAT scale 22 (a, b, c, d =
  .57, .19, .19, .05), 64M undirected edges, 32 Zipf(0.8) labels; cycles
  C4-C8, cliques K4-K8 and induced walks of 8-50 vertices.

The Chung-Lu queries pass the feasibility rule of gpsm_gen_queries (an exact
upper bound on the embedding count, `budget`; slots shrink until it holds).
The R-MAT queries pass the device-exact rule of feasible_queries (the join
finishes within `budget` examined candidates).  Realised sizes are reported
with every result.
"""
from . import _lib
from . import graphs as G

EMBEDDING_BUDGET = 10 ** 8


def feasible_queries(rg, nq, kmin, kmax, kinds, seed, work_budget, per_slot=6, rounds=4, log=None):
    """Queries that pass an exact device-measured feasibility rule.

    For output slot i, per_slot candidates of kind kinds[i % len(kinds)] are
    drawn (sizes uniform in the round's range, no host-side rule), matched in
    one batch with a join work budget, and the first candidate whose join
    finishes within the budget is kept.  Slots left empty are drawn again in
    the next round with the upper size bound halved towards kmin; after
    `rounds` rounds a slot that is still empty is dropped.  Every decision is
    an exact property of the query, so the set does not depend on the device
    or thread count.  Returns (queries, number of dropped slots)."""
    picks = [None] * nq
    hi = kmax
    for r in range(rounds):
        for ki, kind in enumerate(kinds):
            slots = [s for s in range(ki, nq, len(kinds)) if picks[s] is None]
            if not slots:
                continue
            cands = G.queries(rg.g, len(slots) * per_slot, kmin, hi, kind, seed + 7919 * ki + 104729 * r, 0)
            res = rg.match_batch(cands, work_budget=work_budget)
            for si, slot in enumerate(slots):
                for j in range(per_slot):
                    if res[si * per_slot + j].status == 0:
                        picks[slot] = cands[si * per_slot + j]
                        break
            if log:
                log("  feasibility round %d, kind %d, sizes %d-%d: %d of %d slots filled"
                    % (r, kind, kmin, hi, sum(picks[s] is not None for s in slots), len(slots)))
        hi = max(kmin, (kmin + hi) // 2)
    out = [q for q in picks if q is not None]
    return out, nq - len(out)


def er10k_walk6(seed=1):
    return G.er10k_walk6(seed)


def chunglu1m_walk8_16(seed=2, budget=EMBEDDING_BUDGET, nq=100):
    g = G.chunglu_1m(seed)
    return g, G.queries(g, nq, 8, 16, _lib.Q_WALK_HALF, seed, budget)


def chunglu1m_mix4_32(seed=2, budget=EMBEDDING_BUDGET, nq=256, g=None):
    g = g if g is not None else G.chunglu_1m(seed)
    return g, G.queries(g, nq, 4, 32, _lib.Q_WALK_HALF, seed + 1000, budget)


def rmat22_64m(seed=5, scale=22, m=64_000_000):
    g = G.rmat(scale, m, 0.57, 0.19, 0.19, seed)
    return G.set_labels(g, 32, 0.8, seed)


def rmat_shape_queries_device(rg, seed=5, work_budget=EMBEDDING_BUDGET, per_kind=8, per_slot=6, log=None):
    """Cycles C4-C8, cliques K4-K8 and induced walks of 8-50 vertices drawn at
    their sizes and kept by the device-exact rule (feasible_queries).  On the
    R-MAT graph the spanning-tree bound of the host rule passes almost no
    cycle or clique of 4 or more vertices (paths through 10^5-degree hubs),
    and evaluating it for one 8-50-vertex walk takes minutes of host time.
    Returns (queries, dropped slots)."""
    qs, dropped = [], 0
    for kind, kmin, kmax, s in ((_lib.Q_CYCLE, 4, 8, seed), (_lib.Q_CLIQUE, 4, 8, seed + 1),
                                (_lib.Q_WALK_INDUCED, 8, 50, seed + 2)):
        q, d = feasible_queries(rg, per_kind, kmin, kmax, [kind], s, work_budget, per_slot, log=log)
        qs += q
        dropped += d
    return qs, dropped


def rmat_shape_queries(g, seed=5, budget=EMBEDDING_BUDGET, per_kind=8):
    """The same shapes under the host rule (sizes shrink until the bound holds)."""
    qs = []
    qs += G.queries(g, per_kind, 4, 8, _lib.Q_CYCLE, seed, budget)
    qs += G.queries(g, per_kind, 4, 8, _lib.Q_CLIQUE, seed + 1, budget)
    qs += G.queries(g, per_kind, 8, 50, _lib.Q_WALK_INDUCED, seed + 2, budget)
    return qs


def build(name, seed=None, budget=EMBEDDING_BUDGET):
    """(graph, queries) of a named workload."""
    if name == "er10k_walk6":
        return er10k_walk6(*(() if seed is None else (seed,)))
    if name == "chunglu1m_walk8_16":
        return chunglu1m_walk8_16(2 if seed is None else seed, budget)
    if name == "chunglu1m_mix4_32":
        return chunglu1m_mix4_32(2 if seed is None else seed, budget)
    if name == "rmat22_64m":
        g = rmat22_64m(5 if seed is None else seed)
        return g, None  # queries need the device (rmat_shape_queries_device)
    raise ValueError("unknown workload %r" % name)


def run(name, steps=5, autotune=False, plain=True, seed=None, budget=EMBEDDING_BUDGET, log=print):
    """Times one workload on the device: the batched launch with the
    heuristic and with the learned join order (when trained weights exist),
    and the plain per-query path; all three must agree on every count and
    hash.  Returns a dict of rates (queries/s, embeddings/s) and kernel times."""
    import time

    from . import planner
    from .matcher import Context

    t0 = time.time()
    g, qs = build(name, seed, budget)
    log("%s: graph of %d vertices and %d undirected edges built in %.1f s" % (name, g.n, g.adj // 2, time.time() - t0))
    with Context(0) as ctx:
        rg = ctx.upload(g)
        log("graph resident on %s" % ctx.device_name)
        dropped = None
        if qs is None:
            qs, dropped = rmat_shape_queries_device(rg, 5 if seed is None else seed, budget, log=log)
        t_gen = time.time() - t0
        log("%s: n=%d, %d undirected edges, %d queries (sizes %s), generated in %.1f s"
            % (name, g.n, g.adj // 2, len(qs), [q.n for q in qs], t_gen))
        out = {"workload": name, "n": g.n, "undirected_edges": g.adj // 2, "queries": len(qs),
               "query_sizes": {str(k): [q.n for q in qs].count(k) for k in sorted({q.n for q in qs})},
               "query_edges": [q.m for q in qs], "dropped_slots": dropped, "generation_s": t_gen}
        out["device"] = ctx.device_name
        if autotune:
            ms, tune = rg.autotune(qs[:min(len(qs), 64)])
            out["tuning"] = tune
            log("autotuned: %s (%.2f ms)" % (tune, ms))
        modes = [("heuristic", _lib.ORDER_HEURISTIC)]
        if planner.install(ctx):
            modes.append(("learned", _lib.ORDER_LEARNED))
        ref = None
        for label, mode in modes:
            rg.match_batch(qs, order_mode=mode)
            ctx.profile(1)
            t0 = time.perf_counter()
            for _ in range(steps):
                res = rg.match_batch(qs, order_mode=mode)
            dt = (time.perf_counter() - t0) / steps
            ks = ctx.kernel_stats()
            ctx.profile(0)
            got = [(r.count, r.hash, r.status) for r in res]
            assert ref is None or got == ref, "%s order changed a result" % label
            ref = got
            emb = sum(r.count for r in res)
            out["batched_" + label] = {
                "queries_per_s": len(qs) / dt, "embeddings_per_s": emb / dt, "ms_per_batch": 1e3 * dt,
                "join_work": sum(r.join_work for r in res),
                "kernels_ms": {k: v["ms"] / steps for k, v in ks.items()}}
            log("batched %-9s %9.1f queries/s  %.3g embeddings/s  %.2f ms/batch"
                % (label, len(qs) / dt, emb / dt, 1e3 * dt))
        out["embeddings"] = sum(c for c, _, _ in ref)
        out["statuses"] = sorted({s for _, _, s in ref})
        if plain:
            rg.match(qs[0])
            t0 = time.perf_counter()
            rp = [rg.match(q) for q in qs]
            dt = time.perf_counter() - t0
            assert [(r.count, r.hash, r.status) for r in rp] == ref, "plain path changed a result"
            out["plain_heuristic"] = {"queries_per_s": len(qs) / dt, "embeddings_per_s": out["embeddings"] / dt}
            log("plain     heuristic %9.1f queries/s" % (len(qs) / dt))
    return out


def loaders(name, tmpdir, text=True, log=print):
    """Times the ways a data graph reaches the device: parsing the text
    format (the reference's reader, io/IO.cpp:46-100) and uploading the CSR,
    mapping a binary CSR cache and uploading it, and building the CSR on the
    device from the edge list (gpsm_graph_upload_edges)."""
    import ctypes as C
    import os
    import time

    from .matcher import Context

    g, _ = build(name)
    src, dst = g.edges()
    out = {"workload": name, "n": g.n, "undirected_edges": int(len(src))}
    log("%s: %d vertices, %d edges" % (name, g.n, len(src)))
    with Context(0) as ctx:
        ctx.set_hub_policy(512, 0)  # graph arrays only: the same work on every path
        cache = os.path.join(tmpdir, name + ".csr")
        t0 = time.perf_counter()
        G.save_csr(g, cache)
        out["cache_write_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        m = G.load_csr(cache)
        rg = ctx.upload(m)
        out["cache_map_upload_s"] = time.perf_counter() - t0
        rg.free()
        m.close()
        log("binary cache: write %.2f s, map + upload %.2f s" % (out["cache_write_s"], out["cache_map_upload_s"]))
        t0 = time.perf_counter()
        rd = ctx.upload_edges(g.n, src, dst, g.labels)
        out["device_build_s"] = time.perf_counter() - t0
        assert rd.g.adj == g.adj
        rd.free()
        log("device build from the edge list (incl. reading the CSR back): %.2f s" % out["device_build_s"])
        if text:
            path = os.path.join(tmpdir, name + ".g")
            t0 = time.perf_counter()
            check = _lib.lib().gpsm_io_write_graph(path.encode(), 0, 0, g.n, _lib.ptr(g.labels), len(src),
                                                   _lib.ptr(src), _lib.ptr(dst))
            _lib.lib().gpsm_io_write_end(path.encode())
            out["text_write_s"] = time.perf_counter() - t0
            out["text_bytes"] = os.path.getsize(path)
            log("text: written in %.1f s (%d bytes)" % (out["text_write_s"], out["text_bytes"]))
            t0 = time.perf_counter()
            ng, nv, ne = C.c_uint32(), C.c_uint64(), C.c_uint64()
            rc = _lib.lib().gpsm_io_count_graphs(path.encode(), C.byref(ng), C.byref(nv), C.byref(ne))
            out["text_parse_s"] = time.perf_counter() - t0
            assert rc == 0 and check == 0 and ng.value == 1
            log("text: parsed by the driver's reader in %.1f s" % out["text_parse_s"])
            os.remove(path)
        os.remove(cache)
    return out


def main():
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Times a benchmark workload on the device")
    ap.add_argument("workload", choices=["er10k_walk6", "chunglu1m_walk8_16", "chunglu1m_mix4_32", "rmat22_64m"])
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--budget", type=float, default=EMBEDDING_BUDGET)
    ap.add_argument("--json", default=None)
    ap.add_argument("--loaders", metavar="TMPDIR", default=None,
                    help="time text parsing, the binary cache and the device CSR build instead")
    ap.add_argument("--no-text", action="store_true", help="with --loaders: skip the text format")
    args = ap.parse_args()
    if args.loaders:
        out = loaders(args.workload, args.loaders, not args.no_text, log=lambda s: print(s, flush=True))
    else:
        out = run(args.workload, args.steps, args.autotune, not args.no_plain, budget=int(args.budget),
                  log=lambda s: print(s, flush=True))
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
