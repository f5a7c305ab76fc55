"""The profile summarisers on synthetic kernel traces (CPU only): the
per-kernel rows and the whole-operation row of the scan and compaction
summary, the pairing of gradient launches in the planner network's, and the
bench profile's kernel names, counter sums and overlapped join launches."""
import csv
import json

from gpsm_amd import mlp_throughput as MT
from gpsm_amd import scan_throughput as ST


def _trace(path, rows):
    """rows: (kernel name, duration in us[, grid]) in launch order."""
    t = 1_000_000
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X"])
        w.writeheader()
        for r in rows:
            name, us = r[0], r[1]
            w.writerow({"Kernel_Name": "gpsm::(anonymous namespace)::" + name + "(args)", "Start_Timestamp": t,
                        "End_Timestamp": t + int(us * 1000), "Grid_Size_X": r[2] if len(r) > 2 else 0})
            t += int(us * 1000) + 500


def _table(md):
    return [[c.strip() for c in line.strip("|").split("|")] for line in open(md) if line.startswith("| ")]


def test_scan_summary_rows_and_whole_operation(tmp_path):
    n = 1 << 20
    _trace(tmp_path / "t.csv", [("scan_reduce_kernel", 10.0), ("scan_sums_kernel", 2.0),
                                ("scan_tile_kernel", 20.0)] * 2)
    out = str(tmp_path / "s.md")
    ST.summarise(str(tmp_path / "t.csv"), {"mode": "scan", "n": n}, out, "", False)
    rows = {r[1]: r for r in _table(out)[1:]}
    assert set(rows) == {"scan_reduce_kernel", "scan_sums_kernel", "scan_tile_kernel", "all of the above"}
    # reduce: 4n bytes in 10 us; tile: 8n in 20 us; sums: 16 per tile of 4096
    assert float(rows["scan_reduce_kernel"][4]) == round(4 * n / 10e-6 / 1e9)
    assert int(rows["scan_sums_kernel"][3]) == 16 * (n // 4096)
    whole = rows["all of the above"]
    assert float(whole[2]) == 32.0 and int(whole[3]) == 8 * n
    assert abs(float(whole[5]) - 100 * 8 * n / 32e-6 / 1e9 / 8000) < 0.05


def test_compaction_summary_counts_every_kernel(tmp_path):
    n, ids = 1 << 24, 12345
    _trace(tmp_path / "c.csv", [("seg_popcount_kernel", 3.0), ("seg_offsets_kernel", 1.0),
                                ("compact_rows_kernel", 6.0)])
    out = str(tmp_path / "c.md")
    ST.summarise(str(tmp_path / "c.csv"), {"mode": "compact", "n": n, "density": 0.01, "ids": ids}, out, "", False)
    rows = {r[1]: r for r in _table(out)[1:]}
    assert float(rows["all of the above"][2]) == 10.0
    assert int(rows["all of the above"][3]) == n // 8 + 4 * ids + 4 * (n // 256)
    assert int(rows["seg_offsets_kernel"][3]) == 8 * ((n // 32 + 2047) // 2048) + 4


def test_mlp_summary_pairs_each_gradient_with_its_weight_kernel(tmp_path):
    # 256 samples: 8 waves of 64 lanes (grid 512); a gradient is mlp_grad_kernel
    # followed by mlp_wgrad_kernel, timed together
    _trace(tmp_path / "m.csv", [("mlp_forward_kernel", 5.0, 512), ("mlp_grad_kernel", 7.0, 512),
                                ("mlp_wgrad_kernel", 3.0, 256), ("adam_kernel", 1.0, 64),
                                ("mlp_grad_kernel", 9.0, 512), ("mlp_wgrad_kernel", 5.0, 256)])
    run = tmp_path / "run.json"
    run.write_text(json.dumps([{"B": 256, "forward_samples_per_s": 1.0, "grad_samples_per_s": 2.0,
                                "train_step_samples_per_s": 3.0}]))
    out = str(tmp_path / "m.md")
    MT.summarise(str(tmp_path / "m.csv"), str(run), out, "cmd")
    rows = _table(out)
    grad = [r for r in rows if r[0] == "gradient (both kernels)"]
    assert len(grad) == 1 and grad[0][1] == "256" and grad[0][2] == "2" and float(grad[0][3]) == 12.0
    fwd = [r for r in rows if r[0] == "forward"]
    assert fwd[0][1] == "256" and float(fwd[0][3]) == 5.0


def test_profile_summary_helpers(tmp_path):
    from gpsm_amd import profile_summary as PS
    assert PS.short("void gpsm::(anonymous namespace)::join_tree_kernel<true>(gpsm::JoinArgs)") == \
        "join_tree_kernel_counting"
    assert PS.short("void gpsm::(anonymous namespace)::join_tree_kernel<false>(gpsm::JoinArgs)") == "join_tree_kernel"
    assert PS.short("gpsm::(anonymous namespace)::scan_tile_kernel(unsigned int const*)") == "scan_tile_kernel"
    # per-dispatch counters are summed over their rows (one per XCD or SE)
    p = tmp_path / "pmc.csv"
    with open(p, "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for d, k, v in ((1, "a_kernel", 2.0), (1, "a_kernel", 3.0), (2, "a_kernel", 7.0), (3, "b_kernel", 1.0)):
            w.writerow({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": "FETCH_SIZE", "Counter_Value": v})
    c = PS.counters(str(p))
    assert sorted(c["a_kernel"]["FETCH_SIZE"]) == [5.0, 7.0] and c["b_kernel"]["FETCH_SIZE"] == [1.0]
    # the tree and cyclic join launches of a batch overlap: first start to last end
    with open(tmp_path / "j.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for name, b, e in (("join_tree_kernel<false>", 0, 9000), ("join_cyclic_kernel<false>", 1000, 12000),
                           ("join_tree_kernel<true>", 20000, 90000),
                           ("join_tree_kernel<false>", 100000, 104000), ("join_cyclic_kernel<false>", 100500, 103000)):
            w.writerow({"Kernel_Name": name, "Start_Timestamp": b, "End_Timestamp": e})
    assert PS.join_pair_us(str(tmp_path / "j.csv")) == [12.0, 4.0]
    per = {"join_tree_kernel": {"n": 2, "x": 10}, "join_cyclic_kernel": {"n": 2, "x": 4}}
    PS.add_join_total(per, ["x"], "n")
    assert per["join_kernel"] == {"n": 2, "x": 14}
