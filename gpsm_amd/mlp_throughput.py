"""Throughput of the planner network's kernels (32 -> 128 -> 128 -> 1, BF16
MFMA) on one GPU: forward (inference), forward + backward (gradient) and one
training step, at the training mini-batch (256 samples) and at large batches.

    python -m gpsm_amd.mlp_throughput --json out.json
    rocprofv3 --kernel-trace --stats -d DIR -o mlp -- python3 -m gpsm_amd.mlp_throughput
    python -m gpsm_amd.mlp_throughput --summarise DIR/mlp_kernel_trace.csv --run out.json --out FILE.md

The first form times the C-ABI calls with the host clock (inputs copied from
host memory on every call, as the training loop does).  The summary reads a
kernel trace and reports each kernel's samples/s and MFMA rate per batch
size from its own durations.
"""
import argparse
import collections
import csv
import json
import re
import time

import numpy as np

F, H = 32, 128
# MFMA FLOP per sample as the kernels compute them (2 per multiply-add):
# forward: layer 1 (F x H) and layer 2 (H x H); the gradient adds
# gW2 = d2^T h1 (H x H), dh1 = d2 W2 (H x H) and gW1 = d1^T x (H x F).
MFMA_FLOP_FWD = 2 * (F * H + H * H)
MFMA_FLOP_GRAD = MFMA_FLOP_FWD + 2 * (H * H + H * H + H * F)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
TILE = 32                  # samples per wave


def run(sizes, reps):
    from .matcher import Context
    from .planner import MLP
    out = []
    rng = np.random.default_rng(7)
    with Context(0) as ctx:
        m = MLP(ctx, seed=3)
        for B in sizes:
            X = rng.standard_normal((B, F)).astype(np.float32)
            Y = rng.standard_normal(B).astype(np.float32)
            rec = {"B": B}
            for name, fn in (("forward", lambda: m.forward(X)), ("grad", lambda: m.grad(X, Y)),
                             ("train_step", lambda: m.train_step(X, Y, 1e-4))):
                fn()
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn()
                dt = (time.perf_counter() - t0) / reps
                rec[name + "_ms"] = round(dt * 1e3, 4)
                rec[name + "_samples_per_s"] = round(B / dt, 1)
            print(json.dumps(rec), flush=True)
            out.append(rec)
        m.free()
    return out


def summarise(trace_csv, run_json, out_md, command):
    disp = []
    for r in csv.DictReader(open(trace_csv)):
        k = re.search(r"(mlp_\w+_kernel|adam_kernel|refresh_kernel)", r["Kernel_Name"])
        if k:
            disp.append((int(r["Start_Timestamp"]), k.group(1), int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0),
                         (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3))
    disp.sort()
    # forward launches by samples (one 64-lane wave per 32 samples); a
    # gradient is mlp_grad_kernel (one wave per 32 samples) and the
    # mlp_wgrad_kernel launch after it, timed together
    fwd, grad, parts = collections.defaultdict(list), collections.defaultdict(list), collections.defaultdict(list)
    pending = None
    for _, k, grid, us in disp:
        if k == "mlp_forward_kernel":
            fwd[grid // 64 * TILE].append(us)
        elif k == "mlp_grad_kernel":
            pending = (grid // 64 * TILE, us)
            parts[("mlp_grad_kernel", pending[0])].append(us)
        elif k == "mlp_wgrad_kernel" and pending:
            grad[pending[0]].append(pending[1] + us)
            parts[("mlp_wgrad_kernel", pending[0])].append(us)
            pending = None
        else:
            parts[(k, 0)].append(us)
    lines = ["# Planner network kernels: samples/s and MFMA rate", "",
             "Command (from the repository root, one MI355X):", "", "```", command, "```", "",
             "Network 32 -> 128 -> 128 -> 1 (ReLU), BF16 operands, FP32 accumulation. Forward: one wave per %d "
             "samples. Gradient: `mlp_grad_kernel` (one wave per %d samples: forward, backward to the deltas, bias "
             "and W3 gradients) writes the operands of W2's and W1's gradients in 16-sample blocks, and "
             "`mlp_wgrad_kernel` forms those two as split-K GEMMs (at most 64 workgroups of four waves, one flush "
             "of atomics per workgroup). MFMA FLOP per sample: forward %d, forward + backward %d "
             "(v_mfma_f32_32x32x16_bf16). Peak: %.0f TFLOP/s dense BF16."
             % (TILE, TILE, MFMA_FLOP_FWD, MFMA_FLOP_GRAD, PEAK_BF16_TFLOPS), "",
             "Kernel durations from the trace (per call, averaged over the calls of one batch size):", "",
             "| pass | samples per call | calls | average us | samples/s | MFMA TFLOP/s | % of peak |",
             "|---|---|---|---|---|---|---|"]
    for name, table, flop in (("forward", fwd, MFMA_FLOP_FWD), ("gradient (both kernels)", grad, MFMA_FLOP_GRAD)):
        for samples, ds in sorted(table.items()):
            us = sum(ds) / len(ds)
            tf = flop * samples / (us * 1e-6) / 1e12
            lines.append("| %s | %d | %d | %.1f | %.3g | %.3f | %.4f |" % (name, samples, len(ds), us,
                                                                        samples / (us * 1e-6), tf,
                                                                        100 * tf / PEAK_BF16_TFLOPS))
    lines += ["", "| kernel | samples per call | launches | average us |", "|---|---|---|---|"]
    for (k, samples), ds in sorted(parts.items()):
        lines.append("| %s | %s | %d | %.1f |" % (k, samples or "-", len(ds), sum(ds) / len(ds)))
    if run_json:
        lines += ["", "Whole C-ABI calls timed on the host (inputs copied from host memory each call; the training "
                  "loop uses 256-sample mini-batches):", "",
                  "| samples per call | forward samples/s | gradient samples/s | training step samples/s |",
                  "|---|---|---|---|"]
        for rec in json.load(open(run_json)):
            lines.append("| %d | %.3g | %.3g | %.3g |" % (rec["B"], rec["forward_samples_per_s"],
                                                          rec["grad_samples_per_s"], rec["train_step_samples_per_s"]))
    open(out_md, "w").write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="256,4096,65536")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--json")
    ap.add_argument("--summarise", help="kernel trace CSV of a rocprofv3 run of this module")
    ap.add_argument("--run", help="JSON written by --json, for the host-timed table")
    ap.add_argument("--out")
    ap.add_argument("--command", default="")
    a = ap.parse_args()
    if a.summarise:
        summarise(a.summarise, a.run, a.out, a.command)
        return
    recs = run([int(s) for s in a.sizes.split(",")], a.reps)
    if a.json:
        json.dump(recs, open(a.json, "w"), indent=1)


if __name__ == "__main__":
    main()
