"""Host-side parts of bench.py: the counter summary (--summarize-pmc) and its
reading back for the bench line's `traffic` field, on synthetic rocprofv3
counter files (each counter is scaled by the factor that maps the copy
probe's reading onto its known bytes, per-dispatch figures are averaged over
a kernel's dispatches, and a bounce of a launch run as n concurrent chains is
n dispatches); and the refusal of runs on more than one GPU."""
import csv
import json
import os

import bench

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
KERNELS = {"trace": "void (anonymous namespace)::k_trace<false, false>(...)",
           "shade": "void (anonymous namespace)::k_shade<false, 0u>(...)",
           "shadow": "void (anonymous namespace)::k_shadow<false, false>(...)"}


def write_counters(d, counter, copy_kib, per_kernel_kib):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "terrain_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        n = 0
        for v in copy_kib:
            n += 1
            w.writerow({"Dispatch_Id": n, "Kernel_Name": "(anonymous namespace)::k_copy(...)",
                        "Counter_Name": counter, "Counter_Value": v})
        for key, values in per_kernel_kib.items():
            for v in values:
                n += 1
                w.writerow({"Dispatch_Id": n, "Kernel_Name": KERNELS[key], "Counter_Name": counter,
                            "Counter_Value": v})


def test_summary_scales_and_counts_chains(tmp_path, monkeypatch):
    gib_kib = bench.COPY_PROBE_BYTES / 1024.0
    # FETCH reads half the copy's known bytes (factor 2), WRITE all of them (factor 1)
    write_counters(str(tmp_path / "fetch"), "FETCH_SIZE", [gib_kib / 2, gib_kib / 2],
                   {"trace": [100.0, 300.0], "shade": [50.0, 50.0, 50.0, 50.0], "shadow": [10.0, 30.0]})
    write_counters(str(tmp_path / "write"), "WRITE_SIZE", [gib_kib, gib_kib],
                   {"trace": [20.0, 20.0], "shade": [5.0, 15.0, 5.0, 15.0], "shadow": [1.0, 3.0]})
    out = str(tmp_path / "traffic.json")
    bench.summarize_pmc(str(tmp_path / "fetch"), str(tmp_path / "write"), 20, out, bench.TRAFFIC_KERNELS, 4)
    t = json.load(open(out))
    assert abs(t["fetch_copy_factor"] - 2.0) < 1e-12 and abs(t["write_copy_factor"] - 1.0) < 1e-12
    assert t["chains"] == 4 and t["frames_per_launch"] == 20
    tr = t["kernels"]["trace"]
    assert tr["fetch_dispatches"] == 2
    assert abs(tr["fetch_bytes_per_dispatch"] - 200.0 * 1024 * 2) < 1e-6   # mean 200 KiB x 2
    assert abs(tr["write_bytes_per_dispatch"] - 20.0 * 1024) < 1e-6
    assert tr["hbm_bytes_per_dispatch"] == int(200 * 1024 * 2 + 20 * 1024)
    assert tr["hbm_bytes_per_launch"] == tr["hbm_bytes_per_dispatch"] * 4
    sh = t["kernels"]["shade"]
    assert sh["hbm_bytes_per_launch"] == int(50 * 1024 * 2 + 10 * 1024) * 4
    # read back for the bench line: per logical launch, scaled in frames for
    # another launch shape and marked so
    monkeypatch.setattr(bench, "TRAFFIC_FILE", out)
    same, note = bench.measured_traffic(20)
    assert note is None and same["trace"] == tr["hbm_bytes_per_launch"]
    half, note = bench.measured_traffic(10)
    assert note is not None and abs(half["trace"] - tr["hbm_bytes_per_launch"] / 2) < 1e-6


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N with N > 1 exits with a message
    before touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 2
    assert "measures one GPU" in out.stderr and out.stdout == ""
