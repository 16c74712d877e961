"""rocsum (graphics-book-tutorials_amd/rocsum.py), the rocprofv3 CSV summariser, on synthetic CSVs:
kernel-name shortening, per-kernel statistics and counter means, non-numeric values left out (as the
reference's tools/ncusum.cpp skips text-valued metrics), and the copy calibration of FETCH_SIZE and
WRITE_SIZE into device-memory bytes (tests/catch/ncusum.cpp covers the reference's summariser)."""
import csv
import json
import os
import subprocess
import sys

import pytest

from conftest import PKG

sys.path.insert(0, PKG)
import rocsum  # noqa: E402

COUNTER_HEAD = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
                "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
                "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def write_counters(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "x_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(COUNTER_HEAD)
        for i, (k, c, v) in enumerate(rows):
            w.writerow([i, i, "Agent 1", 1, 1, 1, 64, 1, k, 64, 0, 0, 8, 0, 16, c, v, 0, 1])


def write_stats(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "x_kernel_stats.csv"), "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"])
        for k, calls, total, mn, mx in rows:
            w.writerow([k, calls, total, total / calls, 0.0, mn, mx, 0.0])


@pytest.mark.parametrize("full,short", [
    ("wf::gpu::(anonymous namespace)::wf_extend_kernel(wf::SceneView, wf::gpu::PathPool, unsigned int const*)",
     "wf_extend_kernel"),
    ("void wf::gpu::(anonymous namespace)::wf_shade_specular_kernel<2>(wf::SceneView, unsigned int)",
     "wf_shade_specular_kernel<2>"),
    ("void wf::gpu::(anonymous namespace)::scan_reduce_kernel<8, true>(unsigned int const*, int)",
     "scan_reduce_kernel<8, true>"),
    ("__amd_rocclr_fillBufferAligned", "__amd_rocclr_fillBufferAligned"),
])
def test_short_names(full, short):
    assert rocsum.short_name(full) == short


def test_statistics_counters_and_calibration(tmp_path):
    k1 = "void ns::(anonymous namespace)::copy_kernel(float4 const*, float4*, unsigned long)"
    k2 = "ns::(anonymous namespace)::work_kernel(int)"
    write_counters(tmp_path / "cf", [(k1, "FETCH_SIZE", 512.0), (k1, "FETCH_SIZE", 512.0)])
    write_counters(tmp_path / "cw", [(k1, "WRITE_SIZE", 1024.0)])
    write_counters(tmp_path / "f", [(k2, "FETCH_SIZE", 100.0), (k2, "FETCH_SIZE", 300.0), (k2, "FETCH_SIZE", "n/a")])
    write_counters(tmp_path / "w", [(k2, "WRITE_SIZE", 50.0), (k2, "WRITE_SIZE", 50.0)])
    write_stats(tmp_path / "t", [(k2, 4, 4000.0, 900.0, 1100.0), (k1, 1, 10.0, 10.0, 10.0)])
    cal = rocsum.calibrate(str(tmp_path / "cf"), str(tmp_path / "cw"), 1 << 20)
    assert cal["copy_kernel"] == "copy_kernel"
    assert cal["fetch_factor"] == pytest.approx(2.0) and cal["write_factor"] == pytest.approx(1.0)
    doc = rocsum.summarise(str(tmp_path / "t"), [str(tmp_path / "f"), str(tmp_path / "w")], cal)
    w = doc["kernels"]["work_kernel"]
    assert w["calls"] == 4 and w["mean_ns"] == 1000.0 and (w["min_ns"], w["max_ns"]) == (900.0, 1100.0)
    assert w["counters"]["FETCH_SIZE"] == {"dispatches": 2, "total": 400.0, "per_dispatch": 200.0}   # "n/a" left out
    assert w["hbm_bytes_per_dispatch"] == pytest.approx((200.0 * 2.0 + 50.0 * 1.0) * 1024)
    assert w["hbm_gbs"] == pytest.approx(w["hbm_bytes_per_dispatch"] / 1000.0)
    assert doc["summary"]["dominant_kernel"] == "work_kernel"
    assert doc["summary"]["dominant_share"] == pytest.approx(4000.0 / 4010.0)
    md = rocsum.markdown(doc)
    assert md.splitlines()[2].startswith("| `work_kernel` | 4 | 1.0 |")
    out = tmp_path / "s.json"
    r = subprocess.run([sys.executable, os.path.join(PKG, "rocsum.py"), "--trace", str(tmp_path / "t"), "--counters",
                        str(tmp_path / "f"), str(tmp_path / "w"), "--calibrate", str(tmp_path / "cf"),
                        str(tmp_path / "cw"), str(1 << 20), "--json", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert json.load(open(out))["kernels"]["work_kernel"]["calls"] == 4


def test_missing_inputs_are_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        rocsum.kernel_stats(str(tmp_path))
    r = subprocess.run([sys.executable, os.path.join(PKG, "rocsum.py")], capture_output=True, text=True)
    assert r.returncode == 2 and "nothing to summarise" in r.stderr
