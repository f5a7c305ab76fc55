"""Benchmark of the GpSM filter-and-join matcher on one MI355X.

Workload: the batched / learned-planner configuration on a Chung-Lu
power-law data graph (1,000,000 vertices, 8,000,000 undirected edges, gamma
2.1, 16 labels with Zipf(1.0) frequencies): 256 random-walk query slots
(half tree-like, half with every induced edge), each draw taking a size
uniform in 4-32, kept by the device-exact feasibility rule (the exact join
with the reference's heuristic order finishes within 1e10 examined
candidates; resample only, sizes never shrunk).  The set is generated once
by gpsm_amd.workloads (--generate) and read from gpsm_amd/data/.  The
100-slot 8-16-vertex configuration on the same graph keeps no query under
the same rule; its record (draws, lower bounds on the rejected draws'
embedding counts) is reported in the JSON line.

A step is one pass of the matcher over the whole query set (every query's
exact embedding count and set hash), with the data graph resident in HBM.
The headline path is the batched launch (filter supersteps for all queries,
one persistent join grid) with the learned join order when a trained planner
is present, else the reference heuristic order.  Beside it: the heuristic
order on the same path (alternating batches), the same path replayed through
the cached hipGraphs, and the plain path (one query at a time, heuristic
order, the reference's structure).

Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "chunglu1m_mix4_32"
STATED = "chunglu1m_walk8_16"
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic__chunglu1M_mix256.json")
LIMITER_JSON = os.path.join(ROOT, "profiles", "valu__join_chunglu1M_mix256.json")
SAMPLE_MAX_WORK = 5e8  # heaviest query of the CPU sample (examined candidates)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU-baseline sample length (per mode)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tuning (the autotuner otherwise runs before the timed steps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 1 GiB (reference for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--traffic-json", default=TRAFFIC_JSON,
                    help="per-dispatch HBM bytes of each kernel on this workload, from rocprofv3 FETCH_SIZE and "
                         "WRITE_SIZE runs scaled by a copy of known size")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py: the matcher runs on exactly one GPU (the matcher shards by query across GPUs only "
              "in design); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from gpsm_amd import _lib
    from gpsm_amd import planner
    from gpsm_amd import workloads as WL
    from gpsm_amd.matcher import Context

    t0 = time.time()
    qs, rec = WL.load_set(WORKLOAD)
    _, stated = WL.load_set(STATED)
    g = WL.chunglu1m()
    t_gen = time.time() - t0
    sizes = [q.n for q in qs]

    ctx = Context(0)
    rg = ctx.upload(g)
    mode = _lib.ORDER_HEURISTIC
    planner_used = False
    if planner.install(ctx):
        mode = _lib.ORDER_LEARNED
        planner_used = True

    tuning = None
    if not args.no_autotune:
        # filter block, join buffer rows, join workgroups per CU and tasks per
        # workgroup, chosen on the first 64 queries before the timed region
        # (results checked equal)
        _, tuning = rg.autotune(qs[:64], order_mode=mode)
    for _ in range(args.warmup):
        rg.match_batch(qs, order_mode=mode)
    # algorithmic bytes per launch: one untimed step with the device work
    # counters (identical inputs give identical counts every step)
    ctx.profile(2)
    rg.match_batch(qs, order_mode=mode)
    counted = ctx.kernel_stats()
    # timed steps: HIP events around every join launch on the matcher's
    # stream (profile mode 3: no events between the other launches), no counters
    ctx.profile(3)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = rg.match_batch(qs, order_mode=mode)
    t1 = time.perf_counter()
    stats = ctx.kernel_stats()
    # the other kernels' times: one more (untimed) step with events around every launch
    ctx.profile(1)
    rg.match_batch(qs, order_mode=mode)
    for k, v in ctx.kernel_stats().items():
        if k != "join" and not k.startswith("wall:"):
            stats[k] = {"launches": v["launches"] * args.steps, "ms": v["ms"] * args.steps, "bytes": 0.0,
                        "per_step_from": "one untimed step"}
    ctx.profile(0)
    for k, v in stats.items():
        c = counted.get(k)
        v["bytes"] = c["bytes"] / c["launches"] * v["launches"] if c and c["launches"] else 0.0
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    emb = sum(r.count for r in res)
    qps = len(qs) * args.steps / elapsed
    assert all(r.status == 0 for r in res), "a query of the stored set failed"
    want = [(q["count"], int(q["hash"], 16)) for q in rec["queries"]]
    assert [(r.count, r.hash) for r in res] == want, "results differ from the stored set's"

    extra = {}
    # learned vs heuristic order on the same batched path: the two alternate
    # batch by batch, so that drift of the device's clocks affects both alike
    if planner_used:
        cmp = {"learned": [0.0, 0.0], "heuristic": [0.0, 0.0]}
        for _ in range(args.steps):
            for name, m in (("learned", mode), ("heuristic", _lib.ORDER_HEURISTIC)):
                ctx.profile(3)
                t0 = time.perf_counter()
                r = rg.match_batch(qs, order_mode=m)
                cmp[name][0] += time.perf_counter() - t0
                cmp[name][1] += ctx.kernel_stats()["join"]["ms"]
                if name == "heuristic":
                    rh = r
        ctx.profile(0)
        extra["batched_heuristic_qps"] = len(qs) * args.steps / cmp["heuristic"][0]
        extra["learned_over_heuristic"] = cmp["heuristic"][0] / cmp["learned"][0]
        extra["alternating_qps"] = {k: round(len(qs) * args.steps / v[0], 3) for k, v in cmp.items()}
        extra["join_ms_per_step"] = {k: round(v[1] / args.steps, 3) for k, v in cmp.items()}
        assert [(r.count, r.hash) for r in rh] == [(r.count, r.hash) for r in res]
        extra["join_work_learned"] = sum(r.join_work for r in res)
        extra["join_work_heuristic"] = sum(r.join_work for r in rh)
        extra["plans_changed"] = sum(a.order != b.order for a, b in zip(res, rh))
    # the same batched path replayed through the context's cached hipGraphs
    # (captured phases updated in place), alternating with plain launches
    gcmp = {"graph": 0.0, "launches": 0.0}
    i0, u0, r0 = ctx.graph_cache_stats()
    for _ in range(args.steps):
        for name, ug in (("graph", True), ("launches", False)):
            t0 = time.perf_counter()
            r = rg.match_batch(qs, order_mode=mode, use_graph=ug)
            gcmp[name] += time.perf_counter() - t0
            if ug:
                assert [(x.count, x.hash) for x in r] == [(x.count, x.hash) for x in res]
    i1, u1, r1 = ctx.graph_cache_stats()
    extra["graph_replay_qps"] = len(qs) * args.steps / gcmp["graph"]
    extra["graph_over_launches"] = gcmp["launches"] / gcmp["graph"]
    extra["graph_instantiations"] = i1 - i0
    extra["graph_updates"] = u1 - u0
    extra["graph_replays"] = r1 - r0
    if not args.no_plain:
        rg.match(qs[0])
        t0 = time.perf_counter()
        rp = [rg.match(q) for q in qs]
        dt = time.perf_counter() - t0
        extra["plain_heuristic_qps"] = len(qs) / dt
        assert [(r.count, r.hash) for r in rp] == [(r.count, r.hash) for r in res]
        extra["speedup_vs_plain"] = qps / extra["plain_heuristic_qps"]

    # host wall-clock phases of a batch (synchronisations and host work included)
    phases = {k[5:]: round(v["ms"] / args.steps, 3) for k, v in stats.items() if k.startswith("wall:")}
    stats = {k: v for k, v in stats.items() if not k.startswith("wall:")}
    # roofline of the dominant kernel (by time in the timed region)
    dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    name, st = dom
    avg_s = st["ms"] / 1e3 / st["launches"]
    bytes_per_launch = st["bytes"] / st["launches"]
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = None
    if args.traffic_json and os.path.exists(args.traffic_json):
        # HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE runs,
        # scaled by what the same counters read for a copy of known size
        t = json.load(open(args.traffic_json)).get("kernels", {}).get(name + "_kernel")
        if t:
            traffic = round(t["hbm_bytes_per_dispatch"])
    roofline = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": round(bytes_per_launch),
                "launches": st["launches"]}
    if os.path.exists(LIMITER_JSON):
        # what limits the kernel instead, from a rocprofv3 counter run of it
        lim = json.load(open(LIMITER_JSON)).get(name + "_kernel")
        if lim:
            roofline["limiter"] = lim
    kernels = {k: {"ms_per_step": round(v["ms"] / args.steps, 3), "launches_per_step": v["launches"] / args.steps,
                   "GBps": round(v["bytes"] / (v["ms"] / 1e3) / 1e9, 1) if v["bytes"] and v["ms"] else None,
                   "frac_of_hbm_peak": round(v["bytes"] / (v["ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                   if v["bytes"] and v["ms"] else None}
               for k, v in stats.items()}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        # a bounded sample of the same workload: queries in slot order whose
        # device-measured join work is at most SAMPLE_MAX_WORK (a heavier one
        # alone would run for minutes on one core); the rates are scaled to
        # the whole set by join work (the CPU's time is dominated by its join)
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        work = [r.join_work for r in res]
        total_work = sum(work)
        light = [i for i, w in enumerate(work) if w <= SAMPLE_MAX_WORK]
        # one thread: sample queries until the sample length has elapsed
        done, done_work, t0 = 0, 0, time.perf_counter()
        for i in light:
            ro = O.o2_match(g, qs[i], mode=0)
            assert (ro["count"], ro["hash"]) == (res[i].count, res[i].hash), "CPU oracle disagrees with the device"
            done += 1
            done_work += work[i]
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt1 = time.perf_counter() - t0
        rate1 = done_work / dt1  # examined candidates per second, one thread
        # all cores: sample queries up to about the same length
        target = rate1 * threads * args.cpu_seconds
        pick, acc = [], 0
        for i in light:
            if acc >= target:
                break
            pick.append(i)
            acc += work[i]
        t0 = time.perf_counter()
        many = O.o2_match_many(g, [qs[i] for i in pick], threads=threads)
        dt_all = time.perf_counter() - t0
        assert many == [(res[i].count, res[i].hash) for i in pick], "CPU oracle disagrees with the device"
        full_s = dt_all * total_work / max(acc, 1)
        # the independent enumerator (O1, parallel inside each query) beside
        # O2 on the same queries: the lightest by join work, until about a
        # third of the sample length has elapsed; reported, not the baseline
        by_work = sorted(range(len(qs)), key=lambda i: work[i])
        o1_pick, t0 = [], time.perf_counter()
        for i in by_work:
            r1 = O.o1_enumerate(g, qs[i], threads=threads)
            assert (r1["count"], r1["hash"]) == (res[i].count, res[i].hash), "enumerator disagrees with the device"
            o1_pick.append(i)
            if time.perf_counter() - t0 > args.cpu_seconds / 3:
                break
        dt_o1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        o2_same = O.o2_match_many(g, [qs[i] for i in o1_pick], threads=threads)
        dt_o2_same = time.perf_counter() - t0
        assert o2_same == [(res[i].count, res[i].hash) for i in o1_pick], "CPU oracle disagrees with the device"
        cpu = {"value": round(len(qs) / full_s, 4), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": "O2 host restatement (g++ -O3), OpenMP across queries on %d threads, over %d of the %d "
                         "queries (slot order, join work <= %.0e each; %.3g of %.3g examined candidates) in %.1f s, "
                         "scaled to the whole set by join work; every count and hash equal to the device's" % (
                             threads, len(pick), len(qs), SAMPLE_MAX_WORK, acc, total_work, dt_all),
               "single_thread": {"value": round(len(qs) / (dt1 * total_work / max(done_work, 1)), 4), "cores": 1,
                                 "sample": "%d queries (%.3g examined candidates) in %.1f s, scaled by join work"
                                           % (done, done_work, dt1)},
               "enumerator_o1": {"queries": len(o1_pick), "threads": threads, "seconds": round(dt_o1, 2),
                                 "o2_seconds_same_queries": round(dt_o2_same, 2),
                                 "sample": "the %d queries of least join work (%.3g examined candidates), the "
                                           "enumerator parallel inside each query, O2 across them; every count "
                                           "and hash equal to the device's; not the baseline"
                                           % (len(o1_pick), sum(work[i] for i in o1_pick))},
               "embeddings_per_sec": round(emb / full_s), "cpu": cpu_model(), "nproc": os.cpu_count()}

    out = {
        "metric": "queries/sec + embeddings/sec, learned vs plain join order, same GPU, exact",
        "value": round(qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "chunglu-1M-8M-zipf16/256-walk-queries-4to32-feasible", "n": g.n,
                   "undirected_edges": g.adj // 2, "queries": len(qs),
                   "query_sizes": {str(k): sizes.count(k) for k in sorted(set(sizes))},
                   "query_edges_total": sum(q.m for q in qs),
                   "feasibility": {"rule": "exact device join with the heuristic order within the work budget; "
                                           "fresh size per draw, never shrunk",
                                   "work_budget": rec["rule"]["work_budget"], "slots": rec["rule"]["slots"],
                                   "draws": rec["draws"], "dropped_slots": rec["dropped_slots"]},
                   "stated_8_16": {"slots": stated["rule"]["slots"], "draws": stated["draws"],
                                   "kept": stated["kept"],
                                   "rejected_count_lower_bound_quantiles":
                                       stated["rejected_count_lower_bound_quantiles"]},
                   "order": "learned" if planner_used else "heuristic", "tuning": tuning,
                   "device": ctx.device_name},
        "embeddings_per_step": emb,
        "embeddings_per_sec": round(emb * args.steps / elapsed),
        "join_work_per_step": sum(r.join_work for r in res),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "phase_ms_per_step": phases,
        "workload_load_s": round(t_gen, 1),
    }
    out.update({k: (round(v, 3) if isinstance(v, float) else v) for k, v in extra.items()})
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
