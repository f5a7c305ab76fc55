"""Summaries of rocprofv3 runs of bench.py for profiles/ (text only).

Input: the output directories of the runs (kernel trace with --stats;
--pmc FETCH_SIZE and --pmc WRITE_SIZE of the bench and of its
--copy-calibration copy; --pmc with SQ and GRBM counters), and the JSON line
the traced bench run printed.  Output:

* traffic JSON: HBM bytes per dispatch of every kernel, FETCH_SIZE and
  WRITE_SIZE (KiB per dispatch) each scaled by what the same counter reads
  for the 1 GiB device copy (bytes moved / counter value);
* limiter JSON: the join's VALU issue rate from SQ_INSTS_VALU over the CU
  cycles of its dispatches (GRBM_GUI_ACTIVE summed over the XCDs, divided by
  the XCD count, times the CU count), VALU instructions per examined
  candidate and lane, and the instruction mix;
* kernel-trace Markdown table (calls, average and total duration).

    python -m gpsm_amd.profile_summary --prof prof_out --out profiles
"""
import argparse
import collections
import csv
import json
import os
import re

XCDS = 8
CUS = 256


def counters(path):
    """{kernel name: {counter: [per-dispatch values]}} (counter values of one
    dispatch summed over its rows)."""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    for r in csv.DictReader(open(path)):
        key = r["Dispatch_Id"]
        names[key] = r["Kernel_Name"]
        per[key][r["Counter_Name"]] += float(r["Counter_Value"])
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for d, cs in per.items():
        for c, v in cs.items():
            out[names[d]][c].append(v)
    return out


def short(name):
    m = re.search(r"(\w+_kernel)(<(true|false)>)?", name)
    if not m:
        return name.split("(")[0]
    return m.group(1) + ("" if not m.group(2) else ("_counting" if m.group(3) == "true" else ""))


def add_join_total(per_kernel, key, count_key):
    """'join_kernel': the tree-plan and cyclic-plan launches of a batch
    together (one of each per batch), per batch."""
    parts = [per_kernel[k] for k in ("join_tree_kernel", "join_cyclic_kernel") if k in per_kernel]
    if not parts:
        return
    batches = max(p[count_key] for p in parts)
    tot = {count_key: batches}
    for f in key:
        tot[f] = round(sum(p[f] * p[count_key] for p in parts) / batches)
    per_kernel["join_kernel"] = tot


def join_pair_us(trace_csv):
    """Per batch, first start to last end of its join_tree_kernel and
    join_cyclic_kernel launches (counting variants left out), paired in
    launch order; a batch with one of them counts that one."""
    ev = {"join_tree_kernel": [], "join_cyclic_kernel": []}
    for r in csv.DictReader(open(trace_csv)):
        k = short(r["Kernel_Name"])
        if k in ev:
            ev[k].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    t, c = sorted(ev["join_tree_kernel"]), sorted(ev["join_cyclic_kernel"])
    if not t or not c or len(t) != len(c):
        return [(e - b) * 1e-3 for b, e in (t or c)]
    return [(max(a[1], b[1]) - min(a[0], b[0])) * 1e-3 for a, b in zip(t, c)]


def find(d, suffix):
    for f in os.listdir(d):
        if f.endswith(suffix):
            return os.path.join(d, f)
    raise FileNotFoundError(os.path.join(d, "*" + suffix))


def bench_line(log):
    for line in open(log):
        if line.startswith("{\"metric\""):
            return json.loads(line)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--prof", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--tag", default="chunglu1M_mix256")
    ap.add_argument("--command", default="bash gpsm_amd/profile_bench.sh prof_out && "
                    "python3 -m gpsm_amd.profile_summary --prof prof_out --out profiles")
    args = ap.parse_args()
    P = args.prof
    copy_bytes = 1 << 30
    cf = counters(find(os.path.join(P, "cfetch"), "counter_collection.csv"))
    cw = counters(find(os.path.join(P, "cwrite"), "counter_collection.csv"))
    # the copy kernel: the dispatch with the largest counter value
    fscale = copy_bytes / (max(max(v["FETCH_SIZE"]) for v in cf.values()) * 1024)
    wscale = copy_bytes / (max(max(v["WRITE_SIZE"]) for v in cw.values()) * 1024)
    fe = counters(find(os.path.join(P, "fetch"), "counter_collection.csv"))
    wr = counters(find(os.path.join(P, "write"), "counter_collection.csv"))
    traffic = {}
    for name in fe:
        f = fe[name]["FETCH_SIZE"]
        w = wr.get(name, {}).get("WRITE_SIZE", [0.0])
        fb = sum(f) / len(f) * 1024 * fscale
        wb = sum(w) / len(w) * 1024 * wscale
        traffic[short(name)] = {"dispatches": len(f), "fetch_bytes_per_dispatch": round(fb),
                                "write_bytes_per_dispatch": round(wb), "hbm_bytes_per_dispatch": round(fb + wb)}
    add_join_total(traffic, ("fetch_bytes_per_dispatch", "write_bytes_per_dispatch", "hbm_bytes_per_dispatch"),
                   "dispatches")
    bl = bench_line(os.path.join(P, "trace.log"))
    with open(os.path.join(args.out, "traffic__%s.json" % args.tag), "w") as fh:
        json.dump({"about": "HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of "
                            "bench.py, each scaled by the same counter on a 1 GiB device copy "
                            "(bench.py --copy-calibration)",
                   "command": args.command, "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4),
                   "kernels": traffic}, fh, indent=1)
    # SQ counters of the join (timed variant)
    sq = counters(find(os.path.join(P, "sq"), "counter_collection.csv"))
    lim = {}
    joined = collections.defaultdict(list)
    for name, cs in sq.items():
        if short(name) in ("join_tree_kernel", "join_cyclic_kernel"):
            for c, v in cs.items():
                joined[c].append(v)
    if joined:
        # a batch's two join launches together: per-batch sums of the counters
        n = max(len(v) for v in joined["SQ_INSTS_VALU"])
        sq["join_kernel"] = {c: [sum(sum(x) for x in vs) / n] * n for c, vs in joined.items()}
    for name, cs in sq.items():
        k = short(name)
        if k not in ("join_kernel", "filter_step_kernel", "cet_count_kernel", "cet_write_kernel"):
            continue
        n = len(cs["SQ_INSTS_VALU"])
        avg = {c: sum(v) / len(v) for c, v in cs.items()}
        cu_cycles = avg["GRBM_GUI_ACTIVE"] / XCDS * CUS
        e = {"dispatches": n, "valu_insts_per_dispatch": avg["SQ_INSTS_VALU"],
             "salu_insts_per_dispatch": avg["SQ_INSTS_SALU"], "lds_insts_per_dispatch": avg["SQ_INSTS_LDS"],
             "vmem_rd_insts_per_dispatch": avg["SQ_INSTS_VMEM_RD"],
             "valu_issue_frac": round(avg["SQ_INSTS_VALU"] / cu_cycles, 4),
             "salu_issue_frac": round(avg["SQ_INSTS_SALU"] / cu_cycles, 4)}
        lim[k] = e
    if bl and "join_kernel" in lim:
        j = lim["join_kernel"]
        work = bl.get("join_work_per_step")
        # the SQ run's join dispatches mix both orders; their work per dispatch
        # is within 1% of the learned order's
        if work:
            j["valu_lane_insts_per_candidate"] = round(j["valu_insts_per_dispatch"] * 64 / work, 2)
        us = bl["roofline"]["avg_launch_us"]
        top = max(j["valu_issue_frac"], j["salu_issue_frac"])
        j["limiter"] = ("instruction issue" if top > 0.85 else
                        "latency (VALU issue %.0f%%, SALU issue %.0f%% of one per CU per cycle; neither saturated)"
                        % (100 * j["valu_issue_frac"], 100 * j["salu_issue_frac"]))
        j["note"] = ("issue peaks: one wave64 VALU and one SALU instruction per CU per cycle; *_issue_frac = "
                     "SQ_INSTS_* / (GRBM_GUI_ACTIVE / %d XCDs x %d CUs)" % (XCDS, CUS))
        j["bench_avg_launch_us"] = us
    with open(os.path.join(args.out, "valu__join_%s.json" % args.tag), "w") as fh:
        json.dump({"join_kernel": lim.get("join_kernel"), "others": {k: v for k, v in lim.items() if k != "join_kernel"},
                   "command": args.command}, fh, indent=1)
    # kernel trace table
    rows = list(csv.DictReader(open(find(os.path.join(P, "trace"), "kernel_stats.csv"))))
    lines = ["# Kernel trace: %s" % args.tag, "", "Command (from the repository root, one MI355X):", "",
             "```", args.command, "```", "",
             "| kernel | calls | average us | total ms | % |", "|---|---|---|---|---|"]
    for r in rows[:16]:
        lines.append("| %s | %s | %.1f | %.2f | %.2f |" % (short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3,
                                                         float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    pair = join_pair_us(find(os.path.join(P, "trace"), "kernel_trace.csv"))
    if pair:
        lines += ["", "A batch's tree-plan and cyclic-plan join launches run at the same time on two streams, so "
                  "their durations overlap (and the %% column counts both): from their first start to their last "
                  "end they take %.1f us per batch on average (%d batches of the trace)." % (sum(pair) / len(pair),
                                                                                             len(pair))]
    if bl:
        lines += ["", "The traced run's bench line: join average launch %.1f us over its %d timed launches (HIP "
                  "events), %s queries/s." % (bl["roofline"]["avg_launch_us"], bl["roofline"]["launches"], bl["value"])]
    with open(os.path.join(args.out, "kernels__%s.md" % args.tag), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(json.dumps({"traffic": traffic, "limiter": lim}, indent=1))


if __name__ == "__main__":
    main()
