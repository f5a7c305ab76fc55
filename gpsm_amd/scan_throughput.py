"""HBM throughput of the scan and compaction kernels on one GPU, from their
own durations in a rocprofv3 kernel trace.

    rocprofv3 --kernel-trace -d DIR -o scan --output-format csv -- \\
        python3 -m gpsm_amd.scan_throughput --mode scan --log2n 27
    rocprofv3 --kernel-trace -d DIR -o comp --output-format csv -- \\
        python3 -m gpsm_amd.scan_throughput --mode compact --log2n 30 --density 0.05
    python -m gpsm_amd.scan_throughput --summarise DIR/scan_kernel_trace.csv --mode scan --log2n 27 ...

gpsm_scan_u32 (exclusive scan of n u32 values: a reduce pass reading 4n
bytes, one workgroup scanning the tile sums, a tile pass reading 4n and
writing 4n) and gpsm_compact_bitmap (ids of the set bits of an n-bit bitmap:
a popcount pass reading n/8 bytes, then the compaction reading n/8 and
writing 4 bytes per id and 4 per 256-bit block of the rank index).  Inputs
are copied from host memory by the C-ABI calls; the copies are not kernels
and are not counted.
"""
import argparse
import collections
import csv
import json
import re

import numpy as np

PEAK_GBS = 8000.0
REPS = 5


def run(mode, log2n, density, seed=3):
    from .matcher import Context
    n = 1 << log2n
    rng = np.random.default_rng(seed)
    with Context(0) as ctx:
        if mode == "scan":
            a = rng.integers(0, 1 << 16, size=n, dtype=np.uint32)
            for _ in range(REPS):
                out, tot = ctx.scan(a)
            assert tot == int(a.astype(np.uint64).sum())
            return {"mode": mode, "n": n, "reps": REPS}
        words = np.zeros(n // 32, dtype=np.uint32)
        for b in range(32):
            words |= (rng.random(n // 32) < density).astype(np.uint32) << np.uint32(b)
        ones = int(sum(bin(int(x)).count("1") for x in words[:4096])) * (len(words) // 4096)
        for _ in range(REPS):
            ids = ctx.compact(words, n)
        return {"mode": mode, "n": n, "reps": REPS, "density": density, "ids": int(len(ids)), "ids_estimate": ones}


def kernel_bytes(mode, n, ids):
    if mode == "scan":
        return {"scan_reduce_kernel": 4 * n, "scan_sums_kernel": 16 * ((n + 4095) // 4096), "scan_tile_kernel": 8 * n}
    nseg = (n // 32 + 2047) // 2048  # 2048-word segments
    return {"seg_popcount_kernel": n // 8, "seg_offsets_kernel": 8 * nseg + 4,
            "compact_rows_kernel": n // 8 + 4 * ids + 4 * (n // 256)}


def summarise(trace_csv, rec, out_md, command, append):
    dur = collections.defaultdict(list)
    for r in csv.DictReader(open(trace_csv)):
        m = re.search(r"(\w+_kernel)", r["Kernel_Name"])
        if m:
            dur[m.group(1)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    byt = kernel_bytes(rec["mode"], rec["n"], rec.get("ids", 0))
    lines = [] if append else [
        "# Scan and compaction kernels: HBM throughput", "",
        "Each configuration is one rocprofv3 --kernel-trace run (commands below); per kernel the average of its "
        "%d launches, and its algorithmic bytes (each element read or written once) over that time. Peak: %.0f GB/s."
        % (REPS, PEAK_GBS), "",
        "| configuration | kernel | average us | algorithmic bytes | GB/s | % of peak |", "|---|---|---|---|---|---|"]
    cfg = ("scan of 2^%d u32" % (rec["n"].bit_length() - 1) if rec["mode"] == "scan" else
           "compaction of a 2^%d-bit bitmap, density %.2f (%d ids)" % (rec["n"].bit_length() - 1, rec["density"],
                                                                       rec["ids"]))
    for k, b in byt.items():
        ds = dur.get(k)
        if not ds:
            continue
        us = sum(ds) / len(ds)
        gbs = b / (us * 1e-6) / 1e9
        lines.append("| %s | %s | %.1f | %d | %.0f | %.1f |" % (cfg, k, us, b, gbs, 100 * gbs / PEAK_GBS))
    # the whole operation: its kernels one after another, against the bytes
    # of the operation itself (the scan reads and writes each value once;
    # the compaction reads the bitmap once and writes the ids and the index)
    if all(dur.get(k) for k in byt):
        us = sum(sum(dur[k]) / len(dur[k]) for k in byt)
        b = 8 * rec["n"] if rec["mode"] == "scan" else rec["n"] // 8 + 4 * rec.get("ids", 0) + 4 * (rec["n"] // 256)
        gbs = b / (us * 1e-6) / 1e9
        lines.append("| %s | all of the above | %.1f | %d | %.0f | %.1f |" % (cfg, us, b, gbs, 100 * gbs / PEAK_GBS))
    with open(out_md, "a" if append else "w") as fh:
        fh.write("\n".join(lines) + "\n")
    if command:
        with open(out_md + ".commands", "a") as fh:
            fh.write(command + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["scan", "compact"], required=True)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--density", type=float, default=0.05)
    ap.add_argument("--json")
    ap.add_argument("--summarise", help="kernel trace CSV of a rocprofv3 run of this module")
    ap.add_argument("--run", help="JSON written by --json for that run")
    ap.add_argument("--out")
    ap.add_argument("--append", action="store_true")
    ap.add_argument("--command", default="")
    a = ap.parse_args()
    if a.summarise:
        summarise(a.summarise, json.load(open(a.run)), a.out, a.command, a.append)
        return
    rec = run(a.mode, a.log2n, a.density)
    print(json.dumps(rec), flush=True)
    if a.json:
        json.dump(rec, open(a.json, "w"))


if __name__ == "__main__":
    main()
