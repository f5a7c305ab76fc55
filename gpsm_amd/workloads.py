"""The benchmark workloads, built from seeded generators.

* er10k_walk6: ER 10k vertices / 40k edges, 8 uniform labels, one 6-vertex
  walk query (the size the CPU enumerator finishes in milliseconds).
* chunglu1m_mix4_32: Chung-Lu 1M/8M (gamma 2.1), 16 Zipf(1.0) labels, 256
  walk queries (half tree-like, half induced), each drawn with a size uniform
  in 4-32 and kept by the device-exact feasibility rule below.
* chunglu1m_walk8_16: the same graph with 100 walk query slots of 8-16
  vertices under the same rule (on this graph no draw of 7 or more vertices
  passes it, so every slot is dropped; generate() records the evidence).
* rmat22_64m + rmat22_shapes: R-MAT scale 22 (a, b, c, d = .57, .19, .19,
  .05), 64M undirected edges, 32 Zipf(0.8) labels; cycles and cliques of
  4-8 vertices and induced walks of 8-50 vertices under the same rule.

Feasibility rule (device-exact, resample only).  Every candidate query is
drawn with a fresh size uniform in the slot's range (never shrunk) and is
kept if the device's exact join of it, with the reference's heuristic join
order, finishes within `work_budget` examined candidates (first-table rows
plus candidates examined at every level, JoinOut::work); a query over the
budget is abandoned by the join, so a rejected draw costs at most the
budget.  A slot takes the first kept candidate of its draws, in draw order;
a slot with none after `rounds` rounds of `per_round` draws is dropped and
counted.  Every decision is an exact property of the query, so the set does
not depend on the device or the thread count.  The kept queries therefore
follow the drawn size distribution conditioned on feasibility.

Generating a set runs many budgeted joins, so the sets the benchmark and the
size-scale tests use are generated once (``python -m gpsm_amd.workloads
--generate gpsm_amd/data``) and stored as JSON under gpsm_amd/data/, with the
rule's parameters and the outcome of every slot.
"""
import json
import os
import sys

import numpy as np

from . import _lib
from . import graphs as G

EMBEDDING_BUDGET = 10 ** 8
WORK_BUDGET = 10 ** 10  # examined candidates per query (about 20 ms of join on one MI355X)
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
MATCH_CHUNK = 128  # candidates per batch of the rule


def _match_split(rg, qs, work_budget, order_mode):
    """match_batch, halving the batch while the device runs out of memory
    (the candidate-edge tables of a batch of large draws can be tens of GB)."""
    try:
        return rg.match_batch(qs, work_budget=work_budget, order_mode=order_mode)
    except _lib.GpsmError as e:
        if e.code != -2 or len(qs) == 1:
            raise
    h = len(qs) // 2
    return _match_split(rg, qs[:h], work_budget, order_mode) + _match_split(rg, qs[h:], work_budget, order_mode)


def feasible_queries(rg, nq, kmin, kmax, kinds, seed, work_budget, per_round=8, rounds=8, log=None,
                     order_mode=_lib.ORDER_HEURISTIC):
    """Queries kept by the device-exact rule (module docstring).

    Slot i draws candidates of kind kinds[i % len(kinds)], each with a fresh
    size uniform in [kmin, kmax]; per round, every open slot draws per_round
    candidates, all candidates of a round are matched in one batch with the
    work budget, and a slot keeps its first candidate within the budget.
    Returns (queries in slot order, dropped slot count, stats) where stats
    holds the draws made, the kept draws' (slot, round, index) and, for the
    rejected draws, the embeddings already found when the join was
    abandoned (each a lower bound on that draw's embedding count)."""
    picks = [None] * nq
    where = [None] * nq
    draws = 0
    lower_bounds = []
    for r in range(rounds):
        for ki, kind in enumerate(kinds):
            slots = [s for s in range(ki, nq, len(kinds)) if picks[s] is None]
            if not slots:
                continue
            cands = G.queries(rg.g, len(slots) * per_round, kmin, kmax, kind, seed + 7919 * ki + 104729 * r, 0)
            draws += len(cands)
            res = []
            for c0 in range(0, len(cands), MATCH_CHUNK):  # bounded candidate-edge tables per batch
                res += _match_split(rg, cands[c0:c0 + MATCH_CHUNK], work_budget, order_mode)
            for si, slot in enumerate(slots):
                for j in range(per_round):
                    c = si * per_round + j
                    if res[c].status == 0:
                        if picks[slot] is None:
                            picks[slot] = cands[c]
                            where[slot] = (slot, r, j)
                    elif res[c].status == _lib.EBUDGET:
                        lower_bounds.append(res[c].level_rows[cands[c].n - 1])
            if log:
                log("  feasibility round %d, kind %d, sizes %d-%d: %d of %d open slots filled"
                    % (r, kind, kmin, kmax, sum(picks[s] is not None for s in slots), len(slots)))
    out = [q for q in picks if q is not None]
    stats = {"draws": draws, "kept": [w for w in where if w is not None],
             "rejected_count_lower_bounds": lower_bounds}
    return out, nq - len(out), stats


def er10k_walk6(seed=1):
    return G.er10k_walk6(seed)


def chunglu1m(seed=2):
    return G.chunglu_1m(seed)


def rmat22_64m(seed=5, scale=22, m=64_000_000):
    g = G.rmat(scale, m, 0.57, 0.19, 0.19, seed)
    return G.set_labels(g, 32, 0.8, seed)


# The stored query sets: (graph builder, slots, size range, kinds, seed,
# draws per round, rounds).
SETS = {
    "chunglu1m_mix4_32": ("chunglu1m", 256, 4, 32, [_lib.Q_WALK_TREE, _lib.Q_WALK_INDUCED], 1002, 8, 10),
    "chunglu1m_walk8_16": ("chunglu1m", 100, 8, 16, [_lib.Q_WALK_TREE, _lib.Q_WALK_INDUCED], 2, 8, 1),
    "rmat22_cycles4_8": ("rmat22_64m", 16, 4, 8, [_lib.Q_CYCLE], 5, 8, 4),
    "rmat22_cliques4_8": ("rmat22_64m", 16, 4, 8, [_lib.Q_CLIQUE], 6, 8, 4),
    "rmat22_walks8_50": ("rmat22_64m", 16, 8, 50, [_lib.Q_WALK_INDUCED], 7, 8, 1),
}
GRAPHS = {"chunglu1m": chunglu1m, "rmat22_64m": rmat22_64m}


def set_path(name, data_dir=DATA_DIR):
    return os.path.join(data_dir, name + ".json")


def load_set(name, data_dir=DATA_DIR):
    """(queries, record) of a stored query set."""
    with open(set_path(name, data_dir)) as f:
        rec = json.load(f)
    qs = [G.Query(np.array(q["labels"], dtype=np.uint32), np.array(q["src"], dtype=np.uint32),
                  np.array(q["dst"], dtype=np.uint32)) for q in rec["queries"]]
    return qs, rec


def generate(names, data_dir=DATA_DIR, log=print, work_budget=WORK_BUDGET):
    """Runs the device-exact rule for the named sets and stores them."""
    from .matcher import Context

    os.makedirs(data_dir, exist_ok=True)
    with Context(0) as ctx:
        by_graph = {}
        for name in names:
            by_graph.setdefault(SETS[name][0], []).append(name)
        for gname, group in by_graph.items():
            g = GRAPHS[gname]()
            rg = ctx.upload(g)
            log("%s: %d vertices, %d undirected edges resident" % (gname, g.n, g.adj // 2))
            for name in group:
                _, nq, kmin, kmax, kinds, seed, per_round, rounds = SETS[name]
                qs, dropped, st = feasible_queries(rg, nq, kmin, kmax, kinds, seed, work_budget, per_round, rounds,
                                                   log=log)
                res = rg.match_batch(qs) if qs else []
                lb = sorted(st["rejected_count_lower_bounds"])
                rec = {
                    "workload": name, "graph": gname, "n": g.n, "undirected_edges": g.adj // 2,
                    "rule": {"work_budget": work_budget, "slots": nq, "sizes": [kmin, kmax], "kinds": kinds,
                             "seed": seed, "draws_per_round": per_round, "rounds": rounds,
                             "order": "heuristic"},
                    "draws": st["draws"], "kept": len(qs), "dropped_slots": dropped,
                    "rejected_draws": len(lb),
                    "rejected_count_lower_bound_quantiles": (
                        [lb[int(p * (len(lb) - 1))] for p in (0.0, 0.25, 0.5, 0.75, 1.0)] if lb else None),
                    "query_sizes": {str(k): [q.n for q in qs].count(k) for k in sorted({q.n for q in qs})},
                    "queries": [{"slot": w[0], "round": w[1], "draw": w[2], "labels": q.labels.tolist(),
                                 "src": q.src.tolist(), "dst": q.dst.tolist(), "count": r.count,
                                 "hash": "%016x" % r.hash, "work": r.join_work}
                                for q, r, w in zip(qs, res, st["kept"])],
                }
                with open(set_path(name, data_dir), "w") as f:
                    json.dump(rec, f, separators=(",", ":"))
                log("%s: %d draws, %d kept (sizes %s), %d slots dropped" % (
                    name, st["draws"], len(qs), rec["query_sizes"], dropped))
            rg.free()


def refresh(names, data_dir=DATA_DIR, out_dir=None, log=print):
    """Recomputes the stored sets' per-query count, hash and join work on
    the device (heuristic order) without running the rule again, e.g. after
    a change of the hash definition; the rule's record is kept."""
    from .matcher import Context

    out_dir = out_dir or data_dir
    os.makedirs(out_dir, exist_ok=True)
    with Context(0) as ctx:
        by_graph = {}
        for name in names:
            by_graph.setdefault(SETS[name][0], []).append(name)
        for gname, group in by_graph.items():
            g = GRAPHS[gname]()
            rg = ctx.upload(g)
            for name in group:
                qs, rec = load_set(name, data_dir)
                res = rg.match_batch(qs) if qs else []
                for q, r in zip(rec["queries"], res):
                    assert r.status == 0
                    q["count"], q["hash"], q["work"] = r.count, "%016x" % r.hash, r.join_work
                with open(set_path(name, out_dir), "w") as f:
                    json.dump(rec, f, separators=(",", ":"))
                log("%s: %d queries refreshed" % (name, len(qs)))
            rg.free()


def build(name, seed=None, budget=EMBEDDING_BUDGET):
    """(graph, queries) of a named workload; stored sets are read from
    gpsm_amd/data (queries None for a graph without a stored set)."""
    if name == "er10k_walk6":
        return er10k_walk6(*(() if seed is None else (seed,)))
    if name in SETS:
        qs, _ = load_set(name)
        return GRAPHS[SETS[name][0]](), qs
    if name == "rmat22_shapes":
        g = rmat22_64m()
        qs = []
        for part in ("rmat22_cycles4_8", "rmat22_cliques4_8", "rmat22_walks8_50"):
            qs += load_set(part)[0]
        return g, qs
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
        # the last order's batch replayed through the context's cached
        # hipGraphs, alternating with plain launches batch by batch
        label, mode = modes[-1]
        rg.match_batch(qs, order_mode=mode, use_graph=True)
        i0, u0, r0 = ctx.graph_cache_stats()
        tg = {True: 0.0, False: 0.0}
        for _ in range(steps):
            for ug in (True, False):
                t0 = time.perf_counter()
                r = rg.match_batch(qs, order_mode=mode, use_graph=ug)
                tg[ug] += time.perf_counter() - t0
                assert [(x.count, x.hash, x.status) for x in r] == ref, "hipGraph replay changed a result"
        i1, u1, r1 = ctx.graph_cache_stats()
        out["graph_replay"] = {"order": label, "graph_ms_per_batch": 1e3 * tg[True] / steps,
                               "launches_ms_per_batch": 1e3 * tg[False] / steps,
                               "graph_over_launches": tg[False] / tg[True],
                               "instantiations": i1 - i0, "updates": u1 - u0, "replays": r1 - r0}
        log("hipGraph replay %.3f ms/batch, plain launches %.3f ms/batch (%.2fx; %d instantiations, %d updates, "
            "%d relaunches)" % (1e3 * tg[True] / steps, 1e3 * tg[False] / steps, tg[False] / tg[True], i1 - i0,
                                u1 - u0, r1 - r0))
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


PARTS = {"rmat22_shapes": ("rmat22_cycles4_8", "rmat22_cliques4_8", "rmat22_walks8_50"),
         "chunglu1m_mix4_32": ("chunglu1m_mix4_32",)}


def markdown(out, command):
    """A Markdown summary of a run() result: the stored sets' feasibility
    records, rates per path, the hipGraph comparison and kernel times."""
    L = ["# Workload %s" % out["workload"], "", "Command (from the repository root, one MI355X):", "", "```", command,
         "```", "", "Graph: %d vertices, %d undirected edges; %d queries (sizes %s)."
         % (out["n"], out["undirected_edges"], out["queries"], out["query_sizes"]), ""]
    parts = PARTS.get(out["workload"])
    if parts:
        L += ["Stored sets (gpsm_amd/data; a draw is kept when its exact join with the heuristic order finishes "
              "within the work budget; fresh size per draw, never shrunk):", "",
              "| set | slots | draws | kept | sizes kept | dropped slots | embeddings found before abandoning a "
              "rejected draw (min / median / max) |", "|---|---|---|---|---|---|---|"]
        for name in parts:
            _, rec = load_set(name)
            qs = rec.get("rejected_count_lower_bound_quantiles") or []
            qtxt = " / ".join("%.3g" % qs[i] for i in (0, len(qs) // 2, -1)) if qs else "-"
            L.append("| %s | %d | %d | %d | %s | %d | %s |" % (name, rec["rule"]["slots"], rec["draws"], rec["kept"],
                                                              rec["query_sizes"], rec["dropped_slots"], qtxt))
        L.append("")
    L += ["%d embeddings per batch; every path agrees on every count and hash (checked by the runner)."
          % out["embeddings"], "", "| path | queries/s | embeddings/s | ms per batch | examined candidates |",
          "|---|---|---|---|---|"]
    paths = [k for k in ("batched_heuristic", "batched_learned") if k in out]
    for k in paths:
        v = out[k]
        L.append("| %s | %.2f | %.3g | %.1f | %d |" % (k.replace("_", ", "), v["queries_per_s"], v["embeddings_per_s"],
                                                        v["ms_per_batch"], v["join_work"]))
    if "plain_heuristic" in out:
        v = out["plain_heuristic"]
        L.append("| plain per query, heuristic | %.2f | %.3g | | |" % (v["queries_per_s"], v["embeddings_per_s"]))
    gr = out.get("graph_replay")
    if gr:
        L += ["", "Cached hipGraphs (%s order), alternating with plain launches: %.3f against %.3f ms per batch "
              "(%.2fx; %d instantiations, %d updates, %d relaunches without a capture)."
              % (gr["order"], gr["graph_ms_per_batch"], gr["launches_ms_per_batch"], gr["graph_over_launches"],
                 gr["instantiations"], gr["updates"], gr["replays"])]
    if paths:
        L += ["", "Kernel time per batch (ms, HIP events around every launch):", "",
              "| kernel | " + " | ".join(k.split("_", 1)[1] for k in paths) + " |",
              "|---|" + "---|" * len(paths)]
        names = [n for n in out[paths[0]]["kernels_ms"] if not n.startswith("wall:")]
        for n in names:
            L.append("| %s | %s |" % (n, " | ".join("%.3f" % out[k]["kernels_ms"].get(n, 0.0) for k in paths)))
    return "\n".join(L) + "\n"


def main():
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Times a benchmark workload on the device, or generates the "
                                             "stored query sets")
    ap.add_argument("workload", nargs="?", choices=["er10k_walk6", "chunglu1m_mix4_32", "rmat22_shapes"])
    ap.add_argument("--generate", metavar="DIR", default=None,
                    help="run the feasibility rule for the stored query sets (all, or --sets) into DIR")
    ap.add_argument("--refresh", metavar="DIR", default=None,
                    help="recompute the stored sets' counts, hashes and join work into DIR")
    ap.add_argument("--sets", default=",".join(SETS), help="comma-separated sets for --generate / --refresh")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--budget", type=float, default=EMBEDDING_BUDGET)
    ap.add_argument("--json", default=None)
    ap.add_argument("--markdown", default=None, help="write a Markdown summary of the run to this file")
    ap.add_argument("--loaders", metavar="TMPDIR", default=None,
                    help="time text parsing, the binary cache and the device CSR build instead")
    ap.add_argument("--no-text", action="store_true", help="with --loaders: skip the text format")
    args = ap.parse_args()
    if args.generate:
        generate(args.sets.split(","), args.generate, log=lambda s: print(s, flush=True))
        return
    if args.refresh:
        refresh(args.sets.split(","), out_dir=args.refresh, log=lambda s: print(s, flush=True))
        return
    if args.loaders:
        out = loaders(args.workload, args.loaders, not args.no_text, log=lambda s: print(s, flush=True))
    else:
        out = run(args.workload, args.steps, args.autotune, not args.no_plain, budget=int(args.budget),
                  log=lambda s: print(s, flush=True))
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(out, fh, indent=1)
    if args.markdown and not args.loaders:
        with open(args.markdown, "w") as fh:
            fh.write(markdown(out, "python3 -m gpsm_amd.workloads " + " ".join(sys.argv[1:])))


if __name__ == "__main__":
    main()
