"""The bench's contract on the CPU: one GPU only, and the committed HBM traffic record it
reads is consistent with its own counters and calibration."""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_record_is_copy_calibrated():
    rec = json.load(open(os.path.join(ROOT, "profiles", "meander2M_hbm_traffic.json")))
    cal = rec["calibration"]
    kib = 1024.0
    # the copy moves bytes_read / bytes_written; FETCH_SIZE reads half, WRITE_SIZE all
    assert abs(cal["fetch_factor"] * cal["fetch_kib"] * kib / cal["bytes_read"] - 1.0) < 1e-3
    assert abs(cal["write_factor"] * cal["write_kib"] * kib / cal["bytes_written"] - 1.0) < 1e-3
    for name, k in rec["kernels"].items():
        t = cal["fetch_factor"] * k["fetch_kib"] * kib + cal["write_factor"] * k["write_kib"] * kib
        assert abs(t - k["traffic_bytes"]) <= 1.0, name
        # HBM traffic at least the algorithmic bytes, and within 2x of them
        assert 1.0 <= k["traffic_bytes"] / k["algorithmic_bytes"] < 2.0, name
