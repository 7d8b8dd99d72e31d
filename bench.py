"""Throughput of the MLS-MPM substep on one MI355X.

Default workload (--workload batch-learned): 32 independent indentation
cases of uneven size (point counts log-uniform in [100k, 500k], ~8M points
in total; Young's modulus, tensile strength and disc speed drawn per case)
advanced together by the batched engine with FP32 point storage, FP64
arithmetic and the learned return map (BF16 MLP on MFMA, exact FP64 map as
fallback above a yield residual of 1e-4).  A "step" is one substep of every
case.  Untimed set-up: pairs are harvested from an exact FP64 batched run of
the same cases (which is also the force reference), the MLP is trained on the
GPU, and the plain engines are timed on the same cases for comparison: the
single-case engine case by case (the plain MPM), the batched engine with FP64
and FP32 storage and the exact map.

--workload single: the plain FP64 single-case substep at scale (ice block
5 m x 2 m, 1,600,000 points, grid 1200 x 500).

Prints one JSON line: particle-updates/s, the roofline of the dominant
kernel (the fused G2P+P2G pass, timed with HIP events on its stream; the HBM
traffic per launch from the committed rocprofv3 counter summary), and the
CPU baseline (the C++ restatement's timing build, OpenMP, on this host's
cores at 1 thread and at 16, on a bounded sample of every case of the same
workload).
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
# Algorithmic bytes per particle-update (the roofline numerator): the point
# state read and written once per substep, as the reference's algorithm
# keeps it in memory between substeps -- FP64: read pos 16 + vel 16 + Fe 32
# + Bp 32 + Jp 8 + crushed 1 + grain 2 = 107 B, write the same less grain =
# 105 B, 212 B in all; FP32 state: 55 + 53 = 108 B.  A fused pass performs
# one particle-update per point.  Per grid node the fused pass's share of
# the substep is the P2G result written (24 B) and the G2P velocity read
# (16 B); the grid update kernel has the other 40 B.
FUSED_BYTES_PER_POINT = 212
FUSED_BYTES_PER_NODE = 40
SUBSTEP_BYTES_PER_POINT = 212
SUBSTEP_BYTES_PER_NODE = 80
BATCH_F32_BYTES_PER_POINT = 108
BATCH_F64_BYTES_PER_POINT = 212
# The bytes the fused passes actually have to move are fewer: the next G2P
# recomputes v and Bp before any use, so a fused pass neither reads nor
# writes them; Jp_inv is read only by crushed points, and it and the crushed
# flag are written only when they change.  FP64 single case: read pos 16 +
# Fe 32 + crushed 1 + grain 2, write pos 16 + Fe 32 = 99 B; batched FP32
# storage (a 16-byte record of the 48-bit fixed-point position and the
# crushed flag, a 16-byte record of Fe - I, no grain row): read 16 + 16,
# write 16 + 16 = 64 B.
FUSED_MOVED_BYTES_PER_POINT = 99
BATCH_F32_MOVED_BYTES_PER_POINT = 64
# With FP32 point storage the batched gathers read each node's velocity as
# an FP32 pair: 24 B written by the P2G + 8 B read per node.
BATCH_F32_MOVED_BYTES_PER_NODE = 32
METRIC = "particle-updates/s on 1 GPU: batched+learned vs plain MPM at ≤1% force err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_value(name: str, key: str):
    """A per-launch figure of the fused kernel from the committed rocprofv3
    counter summary (profiles/<name>_counters.txt), if any."""
    path = os.path.join(ROOT, "profiles", f"{name}_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def traffic_from_profiles(name: str):
    """HBM bytes per launch of the fused kernel (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    return profile_value(name, "hbm_bytes_per_launch")


def atomics_from_profiles(name: str, avg_launch_s: float):
    """Global atomic requests at L2 per launch (TCC_ATOMIC_sum) and their rate
    at the measured launch duration."""
    n = profile_value(name, "tcc_atomic_requests_per_launch")
    if n is None or not avg_launch_s:
        return None
    return {"l2_atomic_requests_per_launch": n, "l2_atomic_requests_per_s": n / avg_launch_s}


def cpu_info() -> dict:
    """nproc and the CPU model of this host (the GPU box's host when run there)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(states, target_s: float, label: str) -> dict:
    """The oracle's timing build (oracle/liboracle_fast.so: g++ -O3
    -march=x86-64-v3 -fopenmp -ffp-contract=off, bit-identical to the parity
    build) on this host's cores, over every case of the workload: one
    warm-up substep of each case, then as many substeps of each case as fit
    in about target_s / 3 seconds at 1 thread and target_s * 2/3 at all
    threads.  All threads = the 16 host cores that go with one GPU on the
    box (or fewer if the host has fewer)."""
    import ctypes

    import numpy as np

    import oracle
    from icempm_amd import SimParams

    info = cpu_info()
    threads_all = max(1, min(info["nproc"] or 1, 16))

    def copy_state(p0, soa):
        p = SimParams()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(p0), ctypes.sizeof(SimParams))
        return p, np.ascontiguousarray(soa.copy())

    def rate(threads, seconds):
        work = [copy_state(p0, soa) for p0, soa in states]
        grids = [oracle.new_grid(p) for p, _ in work]
        nts = sum(p.nPts for p, _ in work)
        t0 = time.perf_counter()
        for (p, pts), g in zip(work, grids):       # warm-up: one substep of each case
            oracle.run(p, pts, g, np.zeros(2 * p.n_indenter_subdivisions), 1, nthreads=threads, fast=True)
        one = time.perf_counter() - t0
        nsub = max(1, int(seconds / max(one, 1e-9)))
        t0 = time.perf_counter()
        for (p, pts), g in zip(work, grids):
            oracle.run(p, pts, g, np.zeros(2 * p.n_indenter_subdivisions), nsub, nthreads=threads, fast=True)
        el = time.perf_counter() - t0
        return nts * nsub / el, nsub, el

    v1, n1, e1 = rate(1, target_s / 3)
    va, na, ea = rate(threads_all, target_s * 2 / 3)
    return {
        "value": va,
        "unit": "particle-updates/s",
        "cores": threads_all,
        "kind": "port",
        "sample": f"{label}: every case, 1 warm-up substep then {na} substeps each at {threads_all} threads "
                  f"({ea:.1f} s) and {n1} substeps each at 1 thread ({e1:.1f} s); oracle timing build "
                  f"(g++ -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off)",
        "threads_1": v1,
        "threads_all": va,
        "threads_all_count": threads_all,
        "nproc": info["nproc"],
        "cpu_model": info["cpu_model"],
    }


def fail_gpus() -> int:
    print("bench.py: this build measures one GPU only; runs on more than one GPU are not "
          "supported", file=sys.stderr)
    return 2


def run_single(args) -> dict:
    """The plain FP64 single-case engine on one case (default indent-1.6M)."""
    import numpy as np

    from icempm_amd import CASES, MODE_FAST, MODE_ORDERED, Model

    case = CASES[args.case]
    steps, warmup = args.steps, args.warmup
    log(f"bench: {case.name} n_points={case.n_points} grid={case.grid_x}x{case.grid_y} "
        f"mode={args.mode} warmup={warmup} steps={steps}")
    m = Model(MODE_FAST if args.mode == "fast" else MODE_ORDERED)
    if args.pass_kind is not None:
        m.set_pass_kind(args.pass_kind)
    t0 = time.perf_counter()
    m.load_case(case)
    log(f"bench: ingest {time.perf_counter() - t0:.2f} s")
    pts0 = m.points() if not args.no_cpu_baseline else None
    prms0 = m.prms.copy()

    m.run(warmup)
    m.synchronize()
    log("bench: warm-up done")

    m.kernel_timing(args.timing_stride)
    m.synchronize()
    t0 = time.perf_counter()
    hist = m.run(steps)          # ends with a device->host copy (synchronising)
    m.synchronize()
    el = time.perf_counter() - t0
    ktot, klaunch = m.kernel_timing_read()
    m.kernel_timing(0)
    err = m.error_flag()
    if err or not np.isfinite(hist).all():
        raise RuntimeError("bench: run flagged an out-of-grid point or a non-finite force")

    n = case.n_points
    G = case.grid_x * case.grid_y
    value = n * steps / el
    ms_per_step = el / steps * 1e3
    k_avg_s = (ktot / max(klaunch, 1)) / 1e3
    k_bytes = n * FUSED_BYTES_PER_POINT + G * FUSED_BYTES_PER_NODE
    k_moved = n * FUSED_MOVED_BYTES_PER_POINT + G * FUSED_BYTES_PER_NODE
    achieved = k_bytes / k_avg_s / 1e9 if klaunch else None
    sub_bytes = n * SUBSTEP_BYTES_PER_POINT + G * SUBSTEP_BYTES_PER_NODE
    log(f"bench: {ms_per_step * 1e3:.1f} us/substep, fused kernel {k_avg_s * 1e6:.1f} us "
        f"({klaunch} timed launches), {achieved:.0f} GB/s")

    copy_gbs = m.copy_bandwidth(1 << 30, 20)
    log(f"bench: device copy kernel {copy_gbs:.0f} GB/s (read + write)")
    m.close()

    cpu = None
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline (~{args.cpu_seconds:.0f} s)")
        cpu = cpu_baseline([(prms0, pts0)], args.cpu_seconds, case.name)

    traffic = traffic_from_profiles(f"fused_g2p_p2g_{case.name}")
    return {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
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
            "kernel": "k_particles_wave<G2P,P2G> (fused G2P + P2G)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            # the committed counters' HBM bytes over this run's launch time
            "hbm_frac_counters": (traffic / k_avg_s / 1e9 / HBM_PEAK_GBS) if (traffic and k_avg_s) else None,
            "atomics": atomics_from_profiles(f"fused_g2p_p2g_{case.name}", k_avg_s),
            "counter_profile": f"profiles/fused_g2p_p2g_{case.name}_counters.txt",
            "bytes_per_launch": k_bytes,
            "moved_bytes_per_launch": k_moved,
            "moved_gbs": (k_moved / k_avg_s / 1e9) if k_avg_s else None,
            "avg_launch_us": k_avg_s * 1e6,
            "timed_launches": klaunch,
            "copy_kernel_gbs": copy_gbs,
        },
        "substep_effective_gbs": sub_bytes * steps / el / 1e9,
        "cpu_baseline": cpu,
    }


def rel_l2(a, b) -> float:
    import numpy as np
    n = float(np.linalg.norm(b))
    d = float(np.linalg.norm(a - b))
    return d / n if n > 0 else d


def run_batch_learned(args) -> dict:
    """32 uneven cases, batched, FP32 storage, learned return map; the plain
    engines on the same cases for comparison."""
    import numpy as np

    from icempm_amd import MODE_FAST, Model
    from icempm_amd._lib import BATCH_F32, BATCH_F64
    from icempm_amd.batch import batch_cases, build_batch, ingest_cases
    from icempm_amd.learned import train_return_map

    steps, warmup = args.steps, args.warmup
    cases = batch_cases(args.cases, args.n_min, args.n_max, args.seed)
    t0 = time.perf_counter()
    states = ingest_cases(cases)
    n_total = sum(c.n_points for c in cases)
    nodes = sum(c.grid_x * c.grid_y for c in cases)
    log(f"bench: {len(cases)} cases, {n_total} points, {nodes} grid nodes, ingest "
        f"{time.perf_counter() - t0:.1f} s; warmup={warmup} steps={steps}")

    def prepare(b):
        """capture, instantiate and upload the graphs of the warm-up and the
        timed run (the longer first: its buffers serve both)"""
        for n in sorted({warmup, steps} - {0}, reverse=True):
            b.prepare(n)

    def timed(b, label):
        """W untimed warm-up substeps, then the K timed ones (one graph launch
        each, of graphs prepared beforehand, so that the timed launch follows
        the warm-up without host work in between); the force history of the
        timed substeps"""
        if warmup:
            b.run(warmup)
        ms = b.run(steps)
        hist = [b.force_history(c, steps) for c in range(len(cases))]
        if b.error_flag():
            raise RuntimeError(f"bench: {label}: out-of-grid point")
        if not all(np.isfinite(h).all() for h in hist):
            raise RuntimeError(f"bench: {label}: non-finite force")
        pups = n_total * steps / (ms / 1e3)
        log(f"bench: {label}: {ms / steps * 1e3:.1f} us/substep, {pups:.3e} particle-updates/s")
        return pups, ms, hist

    # 1. the exact FP64 batched run from the initial state: the force
    #    reference of the measured run (its substeps W .. W + K) and the
    #    training pairs, harvested over a fixed horizon whatever --steps is
    horizon = max(args.harvest_steps, warmup + steps)
    b = build_batch(cases, BATCH_F64, states)
    b.set_harvest(args.harvest)
    b.run(horizon)
    if b.error_flag():
        raise RuntimeError("bench: harvest run: out-of-grid point")
    ref_hist = [b.force_history(c, warmup + steps)[warmup:] for c in range(len(cases))]
    X, T, offered = b.read_harvest(args.harvest)
    b.close()
    log(f"bench: exact FP64 run over {horizon} substeps: harvested {offered} pairs, training on {X.shape[0]}")

    # 2. the MLP, trained on the GPU
    t0 = time.perf_counter()
    params, norm, losses = train_return_map(X, T, steps=args.train_steps, batch=8192, lr=3e-3, seed=0)
    train_s = time.perf_counter() - t0
    log(f"bench: trained {args.train_steps} Adam steps in {train_s:.1f} s, loss {losses[-1][1]:.3e}")

    # 3. every engine of the comparison set up before any of them is timed:
    #    the measured configuration (batched FP32 + learned map), the batched
    #    engine with the exact map in FP64 and FP32 storage, and the plain
    #    single-case engine loaded with each case.  Each timed run then
    #    follows GPU work by well under a millisecond of host work, not by
    #    the seconds a batch takes to build or free: the GPU clocks down
    #    when idle, and a 20-substep run that follows 7 ms of host work
    #    measured 10-20% slower per substep than one that follows other GPU
    #    work at once.  Every engine compared here is timed that way.
    #    The comparison batches keep their graphs: the tuning below changes
    #    their chunk tables only if the tuner picks another chunk size than
    #    the default.
    b = build_batch(cases, BATCH_F32, states)
    b.set_learned(params, norm, tol=args.tol)
    b64 = build_batch(cases, BATCH_F64, states)
    b32 = build_batch(cases, BATCH_F32, states)
    prepare(b64)
    prepare(b32)
    plain_steps = max(steps, args.plain_min_steps)
    plain = []
    try:
        for c in cases:
            plain.append(Model(MODE_FAST))
            plain[-1].load_case(c)

        # 4. the measured configuration, autotuned (points per chunk and
        #    re-sort cadence, probed from a checkpoint of the initial state,
        #    which the tuner restores)
        tune_rows = []
        if args.autotune_steps > 0:
            t0 = time.perf_counter()
            tune_rows = b.autotune(args.autotune_steps)
            tune_s = time.perf_counter() - t0
        if args.map_branches is not None:
            b.set_map_branches(args.map_branches)
        chunk, resort_every = b.tuning()
        prepare(b)
        # the fused pass alone, timed with events on the batch stream, from
        # the tuned initial state, which is then restored for the timed run
        b.checkpoint()
        b.kernel_timing(args.timing_stride)
        b.run(args.kernel_steps, use_graph=False)
        ktot, klaunch = b.kernel_timing_read()
        b.kernel_timing(0)
        b.restore()
        b.map_stats(reset=True)
        value, ms, hist = timed(b, "batched FP32 learned")
        branches = b.map_branches()
        acc, fb, ins = b.map_stats(reset=True)
        errs = [rel_l2(h, r) for h, r in zip(hist, ref_hist)]
        probe_branch_stats = b.autotune_branch_stats() if args.autotune_steps > 0 else None
        if args.autotune_steps > 0:
            log(f"bench: autotune {tune_s:.1f} s -> chunk {chunk}, re-sort every {resort_every}: " +
                ", ".join(f"({c},{r}) {t * 1e3:.0f} us" for c, r, t in tune_rows))
        if probe_branch_stats:
            ta, tf, sa, sf = probe_branch_stats
            log(f"bench: network proposals passing the gate over the probes: tension {ta}/{ta + tf}, shear "
                f"{sa}/{sa + sf} -> network on for branches {branches} (1 tension, 2 shear)")
        k_avg_s = (ktot / max(klaunch, 1)) / 1e3
        k_bytes = n_total * BATCH_F32_BYTES_PER_POINT + nodes * FUSED_BYTES_PER_NODE
        k_moved = n_total * BATCH_F32_MOVED_BYTES_PER_POINT + nodes * BATCH_F32_MOVED_BYTES_PER_NODE
        achieved = k_bytes / k_avg_s / 1e9 if klaunch else None
        log(f"bench: fused pass {k_avg_s * 1e6:.1f} us ({klaunch} timed launches), "
            f"{achieved:.0f} GB/s algorithmic")

        # 5. the batched engine with the exact map, same setting
        b64.set_tuning(chunk=chunk, resort_every=resort_every)
        f64_pups, _, _ = timed(b64, "batched FP64 exact")
        b32.set_tuning(chunk=chunk, resort_every=resort_every)
        f32_pups, _, f32_hist = timed(b32, "batched FP32 exact")
        errs32 = [rel_l2(h, r) for h, r in zip(f32_hist, ref_hist)]

        # 6. the plain MPM: the single-case FP64 engine, case by case, the
        #    same substep counts (at least 500 timed substeps per case, so
        #    that a short --steps times the engine, not the per-case launch
        #    and synchronisation, as a case run for its whole simulation
        #    would)
        seq_updates, seq_time = 0, 0.0
        for c, m in zip(cases, plain):
            m.run(warmup)
            m.synchronize()
            t0 = time.perf_counter()
            m.run(plain_steps)
            m.synchronize()
            seq_time += time.perf_counter() - t0
            seq_updates += c.n_points * plain_steps
        copy_gbs = plain[0].copy_bandwidth(1 << 30, 20)
    finally:
        for m in plain:
            m.close()
        b64.close()
        b32.close()
        b.close()
    plain_pups = seq_updates / seq_time
    log(f"bench: plain single-case FP64, case by case: {plain_pups:.3e} particle-updates/s; "
        f"copy kernel {copy_gbs:.0f} GB/s")

    cpu = None
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline (~{args.cpu_seconds:.0f} s)")
        cpu = cpu_baseline(states, args.cpu_seconds, f"the {len(cases)} cases of this workload")

    # the committed counter summary of the kernel the timed run launched
    prof = f"batch_particles_f32_{'learned' if branches else 'exact'}_32cases"
    traffic = traffic_from_profiles(prof)
    return {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered lattices, splitmix64; seeded case parameters)",
        "config": {
            "workload": f"batched + learned: {len(cases)} indentation cases, points log-uniform in "
                        f"[{args.n_min}, {args.n_max}] ({n_total} total, 4 per cell, h=0.012 m), "
                        f"E in [5e9, 10e9] Pa, tensile strength in [0.5e6, 2e6] Pa, disc speed in "
                        f"[0.1, 1.0] m/s, dt=2e-6 s; FP32 point storage, FP64 arithmetic, learned "
                        f"return map (BF16 MLP 5-64-64-3 on MFMA, exact FP64 fallback above a yield "
                        f"residual of {args.tol:g})",
            "cases": len(cases),
            "points": n_total,
            "grid_nodes": nodes,
            "storage": "f32 (positions 48-bit fixed point in cell units)",
            "mlp": "bf16 weights/activations, f32 accumulation",
            "points_per_chunk": chunk,
            "resort_every": resort_every,
            "autotune": [{"chunk": c, "resort_every": r, "ms_per_substep": t} for c, r, t in tune_rows],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": (f"k_batch_particles<F32,G2P,P2G,{'learned' if branches else 'exact'}> "
                       "(fused G2P + P2G, all cases)"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            "hbm_frac_counters": (traffic / k_avg_s / 1e9 / HBM_PEAK_GBS) if (traffic and k_avg_s) else None,
            "atomics": atomics_from_profiles(prof, k_avg_s),
            "counter_profile": f"profiles/{prof}_counters.txt",
            "bytes_per_launch": k_bytes,
            "moved_bytes_per_launch": k_moved,
            "moved_gbs": (k_moved / k_avg_s / 1e9) if k_avg_s else None,
            "avg_launch_us": k_avg_s * 1e6,
            "timed_launches": klaunch,
            "copy_kernel_gbs": copy_gbs,
        },
        # the whole substep (grid update and map kernel included) at the
        # per-unit figure for FP32 state: 108 B per point + 40 B per node
        "substep_effective_gbs": k_bytes * steps / (ms / 1e3) / 1e9,
        "comparison": {
            "plain_single_case_f64_pups": plain_pups,
            "plain_steps_per_case": plain_steps,
            "batched_f64_exact_pups": f64_pups,
            "batched_f32_exact_pups": f32_pups,
            "batched_f32_learned_pups": value,
            "speedup_vs_plain": value / plain_pups,
            "speedup_vs_batched_f64": value / f64_pups,
            "force_rel_l2_max_learned_vs_f64_exact": max(errs),
            "force_rel_l2_max_f32_exact_vs_f64_exact": max(errs32),
            # with both branches off the pass is the exact map's, which keeps
            # no counters: every projection takes the exact map
            "learned_accepted": acc,
            "learned_fallback": fb if branches else None,
            "crushed_inside_surface": ins if branches else None,
            "fallback_rate": fb / max(acc + fb, 1) if branches else 1.0,
            "map_branches_network_on": branches,
            "autotune_probe_branch_stats": (dict(zip(("tension_accepted", "tension_fell_back", "shear_accepted",
                                                      "shear_fell_back"), probe_branch_stats))
                                            if probe_branch_stats else None),
            "harvest_substeps": horizon,
            "pairs_harvested": offered,
            "pairs_trained": int(X.shape[0]),
            "train_steps": args.train_steps,
            "train_seconds": train_s,
            "train_loss": losses[-1][1],
        },
        "cpu_baseline": cpu,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", choices=["batch-learned", "single"], default="batch-learned")
    ap.add_argument("--case", default="indent-1.6M", help="single: the case")
    ap.add_argument("--mode", choices=["fast", "ordered"], default="fast", help="single: the mode")
    ap.add_argument("--pass-kind", type=int, default=None,
                    help="single: particle-pass kernel variant (default: the library's)")
    ap.add_argument("--cases", type=int, default=32, help="batch: number of cases")
    ap.add_argument("--n-min", type=int, default=100_000)
    ap.add_argument("--n-max", type=int, default=500_000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tol", type=float, default=1e-4, help="batch: yield-residual tolerance")
    ap.add_argument("--harvest", type=int, default=4_000_000, help="batch: pairs kept for training")
    ap.add_argument("--harvest-steps", type=int, default=3000,
                    help="batch: substeps of the exact FP64 run that harvests the training pairs")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--plain-min-steps", type=int, default=500,
                    help="batch: fewest timed substeps per case of the plain single-case comparison")
    ap.add_argument("--autotune-steps", type=int, default=200,
                    help="batch: substeps per autotuner candidate (0: keep the defaults)")
    ap.add_argument("--map-branches", type=int, choices=[0, 1, 2, 3], default=None,
                    help="batch: branches the network proposes for, 1 tension + 2 shear "
                         "(default: as the autotuner leaves them)")
    ap.add_argument("--kernel-steps", type=int, default=200,
                    help="batch: stream-launched substeps for the fused-pass timing")
    ap.add_argument("--timing-stride", type=int, default=4,
                    help="bracket every k-th launch of the dominant kernel with HIP events")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        return fail_gpus()
    try:
        out = run_single(args) if args.workload == "single" else run_batch_learned(args)
    except RuntimeError as e:
        log(str(e))
        return 1
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
