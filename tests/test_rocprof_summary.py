"""The rocprofv3 summary tool on small hand-written CSV files (CPU only)."""
import csv
import subprocess
import sys


def _write(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def _run(*args):
    r = subprocess.run([sys.executable, "-m", "icempm_amd.rocprof_summary", *args], capture_output=True, text=True,
                       check=True)
    return r.stdout


def test_trace_groups_by_kernel_and_grid(tmp_path):
    p = tmp_path / "kernel_trace.csv"
    fields = ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]
    rows = [
        {"Kernel_Name": "k_pass", "Start_Timestamp": 0, "End_Timestamp": 100_000, "Grid_Size_X": 2048,
         "Grid_Size_Y": 1, "Grid_Size_Z": 1},
        {"Kernel_Name": "k_pass", "Start_Timestamp": 200_000, "End_Timestamp": 300_000, "Grid_Size_X": 2048,
         "Grid_Size_Y": 1, "Grid_Size_Z": 1},
        {"Kernel_Name": "k_pass", "Start_Timestamp": 400_000, "End_Timestamp": 440_000, "Grid_Size_X": 1024,
         "Grid_Size_Y": 1, "Grid_Size_Z": 1},
    ]
    _write(p, fields, rows)
    out = _run("trace", str(p)).splitlines()
    full = [l for l in out if l.split()[-2:] == ["2048", "k_pass"]]
    half = [l for l in out if l.split()[-2:] == ["1024", "k_pass"]]
    assert len(full) == 1 and len(half) == 1
    calls, avg = full[0].split()[:2]
    assert int(calls) == 2 and abs(float(avg) - 100.0) < 1e-9   # 100,000 ns = 100 us
    calls, avg = half[0].split()[:2]
    assert int(calls) == 1 and abs(float(avg) - 40.0) < 1e-9


def test_counters_largest_grid_and_copy_calibration(tmp_path):
    p = tmp_path / "counter_collection.csv"
    fields = ["Dispatch_Id", "Correlation_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
    rows = []
    # the pass over all cases (grid 2048) and over half of them (grid 1024)
    for d, (g, fetch, write) in enumerate([(2048, 1000.0, 500.0), (2048, 1002.0, 502.0), (1024, 400.0, 200.0)]):
        rows.append({"Dispatch_Id": d, "Correlation_Id": d, "Kernel_Name": "k_pass", "Grid_Size": g,
                     "Counter_Name": "FETCH_SIZE", "Counter_Value": fetch})
        rows.append({"Dispatch_Id": d, "Correlation_Id": d, "Kernel_Name": "k_pass", "Grid_Size": g,
                     "Counter_Name": "WRITE_SIZE", "Counter_Value": write})
    # a copy of 1 MiB read and written counts 512 KiB fetched and 1024 KiB written
    rows.append({"Dispatch_Id": 9, "Correlation_Id": 9, "Kernel_Name": "k_copy", "Grid_Size": 64,
                 "Counter_Name": "FETCH_SIZE", "Counter_Value": 512.0})
    rows.append({"Dispatch_Id": 9, "Correlation_Id": 9, "Kernel_Name": "k_copy", "Grid_Size": 64,
                 "Counter_Name": "WRITE_SIZE", "Counter_Value": 1024.0})
    _write(p, fields, rows)
    out = _run("counters", "--kernel", "k_pass", "--largest-grid", "--copy-bytes", str(1 << 20), str(p))
    assert "2 dispatches" in out
    hbm = [l for l in out.splitlines() if l.startswith("hbm_bytes_per_launch:")][0]
    # medians 1001 KiB fetched (x2 calibration) and 501 KiB written (x1)
    assert float(hbm.split()[1]) == (1001.0 * 2 + 501.0) * 1024
    out_all = _run("counters", "--kernel", "k_pass", "--copy-bytes", str(1 << 20), str(p))
    assert "3 dispatches" in out_all
