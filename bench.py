"""Throughput of the MLS-MPM substep on one MI355X.

Workload: the plain FP64 indentation case at scale -- an ice block 5 m x 2 m
on a jittered lattice with 4 points per cell at h = 0.005 m (1,600,000
points), grid 1200 x 500, rigid disc r = 0.3 m at 0.5 m/s, dt = 1e-6 s.
A "step" is one substep (P2G -> grid update -> G2P with the NACC /
Drucker-Prager return map) over all points; inputs are resident in HBM
before the timed region starts.

Prints one JSON line: particle-updates/s, the roofline of the dominant
kernel (the fused G2P+P2G pass, timed with HIP events on its stream), and the
CPU baseline (the C++ restatement, OpenMP, on this host's cores, on a
bounded prefix of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of the fused G2P+P2G pass (FP64):
#   per point: read pos 16 + Fe 32 + Jp 8 + crushed 1 + grain 2 = 59 B,
#              write pos 16 + vel 16 + Fe 32 + Bp 32 + Jp 8 + crushed 1 = 105 B
#   per grid node: read velocity 16 B (G2P) + add mass/momentum 24 B (P2G)
FUSED_BYTES_PER_POINT = 164
FUSED_BYTES_PER_NODE = 40
# whole-substep algorithmic bytes (point state read + written once: 212 B;
# grid: P2G out 24 + update in 24 / out 16 + G2P in 16 = 80 B per node)
SUBSTEP_BYTES_PER_POINT = 212
SUBSTEP_BYTES_PER_NODE = 80


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(case_name: str):
    """HBM bytes per launch of the fused kernel, from the committed rocprofv3
    counter summary (profiles/fused_g2p_p2g_<case>_counters.txt), if any."""
    path = os.path.join(ROOT, "profiles", f"fused_g2p_p2g_{case_name}_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def cpu_baseline(case, pts0, prms0, target_s: float, threads: int):
    import ctypes

    import numpy as np

    import oracle
    from icempm_amd import SimParams

    p = SimParams()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(prms0), ctypes.sizeof(SimParams))
    pts = np.ascontiguousarray(pts0.copy())
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    oracle.run(p, pts, grid, bins, 1, nthreads=threads)          # warm-up
    t0 = time.perf_counter()
    oracle.run(p, pts, grid, bins, 1, nthreads=threads)
    one = time.perf_counter() - t0
    nsub = max(1, min(200, int(target_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    _, err = oracle.run(p, pts, grid, bins, nsub, nthreads=threads)
    el = time.perf_counter() - t0
    return {
        "value": case.n_points * nsub / el,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{case.name}: {nsub} substeps after 2 warm-up substeps, OpenMP {threads} threads, "
                  f"{el:.1f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--case", default="indent-1.6M")
    ap.add_argument("--mode", choices=["fast", "ordered"], default="fast")
    ap.add_argument("--pass-kind", type=int, default=None,
                    help="particle-pass kernel variant (default: the library's)")
    ap.add_argument("--timing-stride", type=int, default=8,
                    help="bracket every k-th launch of the dominant kernel with HIP events")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the substep of one case on one "
              "MI355X); runs on more than one GPU are not supported", file=sys.stderr)
        return 2

    import numpy as np

    from icempm_amd import CASES, MODE_FAST, MODE_ORDERED, Model

    case = CASES[args.case]
    log(f"bench: {case.name} n_points={case.n_points} grid={case.grid_x}x{case.grid_y} "
        f"mode={args.mode} warmup={args.warmup} steps={args.steps}")
    m = Model(MODE_FAST if args.mode == "fast" else MODE_ORDERED)
    if args.pass_kind is not None:
        m.set_pass_kind(args.pass_kind)
    t0 = time.perf_counter()
    m.load_case(case)
    log(f"bench: ingest {time.perf_counter() - t0:.2f} s")
    pts0 = m.points() if not args.no_cpu_baseline else None
    prms0 = m.prms.copy()

    m.run(args.warmup)
    m.synchronize()
    log("bench: warm-up done")

    m.kernel_timing(args.timing_stride)
    m.synchronize()
    t0 = time.perf_counter()
    hist = m.run(args.steps)          # ends with a device->host copy (synchronising)
    m.synchronize()
    el = time.perf_counter() - t0
    ktot, klaunch = m.kernel_timing_read()
    m.kernel_timing(0)
    err = m.error_flag()
    if err or not np.isfinite(hist).all():
        log("bench: run flagged an out-of-grid point or a non-finite force")
        return 1

    n = case.n_points
    G = case.grid_x * case.grid_y
    value = n * args.steps / el
    ms_per_step = el / args.steps * 1e3
    k_avg_s = (ktot / max(klaunch, 1)) / 1e3
    k_bytes = n * FUSED_BYTES_PER_POINT + G * FUSED_BYTES_PER_NODE
    achieved = k_bytes / k_avg_s / 1e9 if klaunch else None
    sub_bytes = n * SUBSTEP_BYTES_PER_POINT + G * SUBSTEP_BYTES_PER_NODE
    log(f"bench: {ms_per_step * 1e3:.1f} us/substep, fused kernel {k_avg_s * 1e6:.1f} us "
        f"({klaunch} timed launches), {achieved:.0f} GB/s")

    copy_gbs = m.copy_bandwidth(1 << 30, 20)
    log(f"bench: device copy kernel {copy_gbs:.0f} GB/s (read + write)")

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        log(f"bench: CPU baseline ({threads} threads, ~{args.cpu_seconds:.0f} s)")
        cpu = cpu_baseline(case, pts0, prms0, args.cpu_seconds, threads)

    out = {
        "metric": "particle-updates/s on 1 GPU: batched+learned vs plain MPM at ≤1% force err",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered lattice, splitmix64)",
        "config": {
            "workload": f"plain MPM FP64 substep, {case.name}: ice block {case.ncx * case.h:g} m x "
                        f"{case.ncy * case.h:g} m, {n} points (4 per cell, h={case.h:g} m), grid "
                        f"{case.grid_x}x{case.grid_y}, disc r={case.ind_diameter / 2:g} m at "
                        f"{case.ind_velocity:g} m/s, dt={case.dt:g} s",
            "mode": args.mode,
            "points": n,
            "grid_nodes": G,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_particles<G2P,P2G> (fused G2P + P2G)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic_from_profiles(case.name),
            "bytes_per_launch": k_bytes,
            "avg_launch_us": k_avg_s * 1e6,
            "timed_launches": klaunch,
            "copy_kernel_gbs": copy_gbs,
        },
        "substep_effective_gbs": sub_bytes * args.steps / el / 1e9,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
