"""Summaries of rocprofv3 output for the measurement files under profiles/.

  python -m icempm_amd.rocprof_summary stats KERNEL_STATS_CSV [--top N]
      the per-kernel table of a `rocprofv3 --kernel-trace --stats` run
      (calls, average and total duration, share of GPU time)

  python -m icempm_amd.rocprof_summary trace KERNEL_TRACE_CSV [--top N]
      the same per kernel and grid size, from the dispatch list of a
      `rocprofv3 --kernel-trace` run (a kernel launched over parts of its
      work, such as a batched pass over each half of the cases, shows one
      line per launch size)

  python -m icempm_amd.rocprof_summary counters --kernel SUBSTR [--largest-grid] CSV [CSV ...]
      medians per dispatch of every counter in `rocprofv3 --pmc` output files
      for the kernels whose name contains SUBSTR (with --largest-grid, only
      their dispatches of the largest grid size).  When FETCH_SIZE and
      WRITE_SIZE were collected and the run also launched the plain copy
      kernel k_copy over --copy-bytes (read and written per launch), the two
      are calibrated against it (true bytes per counted KiB) and the HBM bytes
      per launch are printed as `hbm_bytes_per_launch: N`, the line bench.py
      reads.

Only rocprofv3's own CSV columns are read (Kernel_Name, Counter_Name,
Counter_Value, Grid_Size; Name, Calls, AverageNs, TotalDurationNs,
Percentage; Start_Timestamp, End_Timestamp, Grid_Size_X/Y/Z).
"""
from __future__ import annotations

import argparse
import collections
import csv
import statistics
import sys


def _rows(paths):
    for p in paths:
        with open(p, newline="") as f:
            yield from csv.DictReader(f)


def _grid(r) -> int:
    """a dispatch's grid size in work-items"""
    if r.get("Grid_Size"):
        return int(float(r["Grid_Size"]))
    g = 1
    for ax in ("X", "Y", "Z"):
        v = r.get(f"Grid_Size_{ax}")
        if v:
            g *= int(float(v))
    return g


def counter_medians(paths, substr: str, largest_grid: bool = False) -> tuple[dict, dict, int]:
    """(medians for the matching kernel, medians for k_copy, matching dispatches)"""
    kv = collections.defaultdict(list)
    cv = collections.defaultdict(list)
    disp = set()
    rows = list(_rows(paths))
    top = max((_grid(r) for r in rows if substr in r.get("Kernel_Name", "")), default=0)
    for r in rows:
        name = r.get("Kernel_Name", "")
        val = float(r["Counter_Value"])
        if substr in name:
            if largest_grid and _grid(r) != top:
                continue
            kv[r["Counter_Name"]].append(val)
            disp.add((r.get("Dispatch_Id"), r.get("Correlation_Id")))
        elif "k_copy" in name:
            cv[r["Counter_Name"]].append(val)
    med = {k: statistics.median(v) for k, v in kv.items()}
    cmed = {k: statistics.median(v) for k, v in cv.items()}
    return med, cmed, len(disp)


def hbm_bytes(med: dict, cmed: dict, copy_bytes: float) -> tuple[float, float, float] | None:
    """(fetch bytes, write bytes, total) per launch, calibrated on k_copy"""
    if not all(k in med for k in ("FETCH_SIZE", "WRITE_SIZE")):
        return None
    if not all(k in cmed and cmed[k] > 0 for k in ("FETCH_SIZE", "WRITE_SIZE")):
        return None
    ff = copy_bytes / (cmed["FETCH_SIZE"] * 1024.0)
    fw = copy_bytes / (cmed["WRITE_SIZE"] * 1024.0)
    f = med["FETCH_SIZE"] * 1024.0 * ff
    w = med["WRITE_SIZE"] * 1024.0 * fw
    return f, w, f + w


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m icempm_amd.rocprof_summary")
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=12)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--top", type=int, default=12)
    c = sub.add_parser("counters")
    c.add_argument("--kernel", required=True)
    c.add_argument("--largest-grid", action="store_true")
    c.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    c.add_argument("csv", nargs="+")
    a = ap.parse_args(argv)
    if a.mode == "stats":
        rows = list(_rows([a.csv]))
        rows.sort(key=lambda r: -float(r.get("TotalDurationNs", 0)))
        print(f"{'calls':>7} {'avg us':>10} {'total ms':>10} {'%':>6}  kernel")
        for r in rows[:a.top]:
            print(f"{int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:10.2f} "
                  f"{float(r['TotalDurationNs']) / 1e6:10.2f} {float(r['Percentage']):6.2f}  {r['Name']}")
        return 0
    if a.mode == "trace":
        groups = collections.defaultdict(list)
        for r in _rows([a.csv]):
            groups[(r["Kernel_Name"], _grid(r))].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        total = sum(sum(v) for v in groups.values()) or 1
        print(f"{'calls':>7} {'avg us':>10} {'total ms':>10} {'%':>6} {'grid':>9}  kernel")
        for (name, g), v in sorted(groups.items(), key=lambda kv: -sum(kv[1]))[:a.top]:
            print(f"{len(v):7d} {sum(v) / len(v) / 1e3:10.2f} {sum(v) / 1e6:10.2f} "
                  f"{100.0 * sum(v) / total:6.2f} {g:9d}  {name}")
        return 0
    med, cmed, n = counter_medians(a.csv, a.kernel, a.largest_grid)
    if not med:
        print(f"no dispatch of a kernel matching {a.kernel!r}", file=sys.stderr)
        return 1
    print(f"kernel matching {a.kernel!r}: {n} dispatches; medians per dispatch")
    for k in sorted(med):
        print(f"  {k:36s} {med[k]:18.1f}")
    hb = hbm_bytes(med, cmed, a.copy_bytes)
    if hb:
        print(f"  k_copy calibration ({a.copy_bytes:.0f} B read and written per launch): FETCH_SIZE "
              f"{cmed['FETCH_SIZE']:.1f} KiB, WRITE_SIZE {cmed['WRITE_SIZE']:.1f} KiB")
        print(f"  HBM bytes per launch: fetched {hb[0]:.0f}, written {hb[1]:.0f}")
        print(f"hbm_bytes_per_launch: {hb[2]:.0f}")
    if "TCC_ATOMIC_sum" in med:
        print(f"tcc_atomic_requests_per_launch: {med['TCC_ATOMIC_sum']:.0f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
