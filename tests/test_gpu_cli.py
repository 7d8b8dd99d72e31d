"""The cm frontend (cli/main.cpp's behaviour): a JSON configuration and an
HDF5 raw point file in, one HDF5 snapshot per output cycle out; in ORDERED
mode the snapshot equals the CPU oracle's state bit for bit."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest
from h5util import H5DUMP, binary

import oracle
from icempm_amd import INDENT_TINY, SimParams, _lib, configure_params, lattice_points
from icempm_amd.build import CLI_BIN
from icempm_amd.model import write_raw_points

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(CLI_BIN) and os.path.exists(H5DUMP)),
                                 reason="cm or h5dump not built / available")]


def test_cm_writes_oracle_snapshots(gpu_available, tmp_path):
    print("cli: cm, 2 cycles ORDERED", flush=True)
    case = INDENT_TINY
    xy, g = lattice_points(case)
    raw = tmp_path / "raw.h5"
    write_raw_points(str(raw), xy, g, case.volume, 7)
    p = SimParams()
    configure_params(p, case)
    cfg = {
        "InitialTimeStep": case.dt, "YoungsModulus": case.youngs, "GridX": case.grid_x,
        "GridY": case.grid_y, "GridXDimension": p.GridXDimension, "SimulationEndTime": 1.0,
        "PoissonsRatio": case.poisson, "Gravity": case.gravity, "Density": case.density,
        "IndDiameter": case.ind_diameter, "IndVelocity": case.ind_velocity, "IndDepth": case.ind_depth,
        "IceCompressiveStrength": case.compressive, "IceTensileStrength": case.tensile,
        "IceShearStrength": case.shear, "DP_phi": case.dp_phi_deg, "DP_threshold_p": case.dp_threshold,
        "GrainVariability": case.grain_variability, "InputRawPoints": str(raw),
    }
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    out = tmp_path / "snap"
    r = subprocess.run([CLI_BIN, str(tmp_path / "c.json"), "--snapshots", str(out), "--ordered", "--pq",
                        "--cycles", "2"], stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=180)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "cm done" in r.stdout
    for f in ("snapshots/d00001.h5", "snapshots/d00002.h5", "pq/pq00001.h5", "pq/pq00002.h5"):
        assert (out / f).exists(), f

    # the oracle: the same ingest and 2 cycles of Model::Step
    snap = out / "snapshots" / "d00002.h5"
    prm = SimParams.from_buffer_copy(binary(snap, "/Params", "u1", tmp_path).tobytes())
    pts = oracle.ingest(p, xy, g, case.volume, prm.nPtsPitch)
    p.nGridPitch = prm.nGridPitch
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    N = p.UpdateEveryNthStep
    _, err = oracle.run(p, pts, grid, bins, 2 * N)
    assert err == 0
    assert prm.SimulationStep == p.SimulationStep and prm.SimulationTime == p.SimulationTime
    dev = binary(snap, "/Points", "<f8", tmp_path).reshape(16, -1)
    assert np.array_equal(dev.view(np.uint64), pts.view(np.uint64))
    assert np.array_equal(binary(snap, "/Indenter_2D", "<f8", tmp_path), bins)


def case_config(case, raw):
    p = SimParams()
    configure_params(p, case)
    return p, {
        "InitialTimeStep": case.dt, "YoungsModulus": case.youngs, "GridX": case.grid_x,
        "GridY": case.grid_y, "GridXDimension": p.GridXDimension, "SimulationEndTime": 1.0,
        "PoissonsRatio": case.poisson, "Gravity": case.gravity, "Density": case.density,
        "IndDiameter": case.ind_diameter, "IndVelocity": case.ind_velocity, "IndDepth": case.ind_depth,
        "IceCompressiveStrength": case.compressive, "IceTensileStrength": case.tensile,
        "IceShearStrength": case.shear, "DP_phi": case.dp_phi_deg, "DP_threshold_p": case.dp_threshold,
        "GrainVariability": case.grain_variability, "InputRawPoints": str(raw),
    }


def run_cm_batch(tmp_path, cases, f32, precision=None, grain_seed=None):
    """cm --batch over the cases, one cycle; with grain_seed, every case but
    the last gets seeded grain ids in [0, 1000)."""
    names, refs = [], []
    rng = np.random.default_rng(grain_seed) if grain_seed is not None else None
    for k, case in enumerate(cases):
        xy, g = lattice_points(case)
        if rng is not None and k < len(cases) - 1:
            g = rng.integers(0, 1000, size=g.shape[0]).astype(np.int16)
        raw = tmp_path / f"raw{k}.h5"
        write_raw_points(str(raw), xy, g, case.volume, 7)
        p, cfg = case_config(case, raw)
        (tmp_path / f"c{k}.json").write_text(json.dumps(cfg))
        names.append(f"c{k}.json")
        refs.append((p, xy, g, case.volume))
    (tmp_path / "m.json").write_text(json.dumps({"cases": names,
                                                 "precision": precision or ("f32" if f32 else "f64")}))
    out = tmp_path / "out"
    r = subprocess.run([CLI_BIN, "--batch", str(tmp_path / "m.json"), "--out", str(out), "--cycles", "1"],
                       stdin=subprocess.DEVNULL, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=180)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "cm done" in r.stdout
    N = refs[0][0].UpdateEveryNthStep
    summary = np.loadtxt(out / "batch.csv", delimiter=",", skiprows=1, ndmin=2)
    assert summary.shape == (1, 4) and summary[0, 1] == N and summary[0, 3] > 0
    hists = []
    for k in range(len(cases)):
        rows = np.loadtxt(out / f"case_{k}.csv", delimiter=",", skiprows=1)
        assert rows.shape == (N, 4) and np.array_equal(rows[:, 0], np.arange(1, N + 1))
        hists.append(rows[:, 2:])
    return refs, N, hists


def oracle_runs(refs, N):
    """The oracle's ingest and N substeps of each case alone: (force history,
    force bins, points, parameters after the run) per case."""
    out = []
    for p, xy, g, vol in refs:
        pts = oracle.ingest(p, xy, g, vol, (xy.shape[0] + 63) // 64 * 64)
        p.nGridPitch = (p.GridTotal + 63) // 64 * 64
        grid = oracle.new_grid(p)
        bins = np.zeros(2 * p.n_indenter_subdivisions)
        href, err = oracle.run(p, pts, grid, bins, N)
        assert err == 0 and np.abs(href).sum() > 0
        out.append((href, bins, pts, p))
    return out


def check_case_snapshots(out, runs, N, tol_bins, tol_x, tmp_path):
    """cm --batch's per-case snapshot after the first cycle against the oracle:
    frame d00001.h5, SimulationStep N and SimulationTime as Model::Step sums
    them, the cycle's force bins (Indenter_2D) and the points' positions."""
    for k, (href, bins, pts, p) in enumerate(runs):
        snap = out / f"case_{k}" / "snapshots" / "d00001.h5"
        assert snap.exists(), snap
        prm = SimParams.from_buffer_copy(binary(snap, "/Params", "u1", tmp_path).tobytes())
        assert prm.SimulationStep == N and prm.SimulationTime == p.SimulationTime
        assert prm.indenter_x == prm.indenter_x_initial + prm.SimulationTime * prm.IndVelocity
        ind = binary(snap, "/Indenter_2D", "<f8", tmp_path)
        assert ind.shape == bins.shape and np.abs(bins).sum() > 0
        e = np.linalg.norm(ind - bins) / np.linalg.norm(bins)
        dev = binary(snap, "/Points", "<f8", tmp_path).reshape(16, -1)
        n = p.nPts
        dx = np.abs(dev[4:6, :n] - pts[4:6, :n]).max()
        print(f"case {k}: snapshot force bins rel L2 {e:.2e}, positions {dx:.2e} m", flush=True)
        assert e <= tol_bins and dx <= tol_x, (k, e, dx)


def test_cm_batch_manifest_matches_oracle(gpu_available, tmp_path):
    """cm --batch: configurations listed in a manifest run as one batch (FP64
    storage); each case's per-substep indenter force over one output cycle
    (case_K.csv) and its snapshot's force bins equal the oracle run of that
    case alone within 1e-6 relative L2 (atomic summation order; measured
    ~2e-12), its positions within 1e-12 m (measured ~2e-15)."""
    from icempm_amd.batch import batch_cases
    print("cli: cm --batch, 3 cases, 1 cycle f64", flush=True)
    refs, N, hists = run_cm_batch(tmp_path, batch_cases(3, 1500, 4000, seed=21), False)
    runs = oracle_runs(refs, N)
    for k, (r, h) in enumerate(zip(runs, hists)):
        href = r[0]
        e = np.linalg.norm(h - href) / np.linalg.norm(href)
        print(f"case {k}: force rel L2 {e:.2e}", flush=True)
        assert e <= 1e-6, (k, e)
    check_case_snapshots(tmp_path / "out", runs, N, 1e-6, 1e-12, tmp_path)


def test_cm_batch_manifest_f32(gpu_available, tmp_path):
    """cm --batch with FP32 storage: each case's force over the 2,500-substep
    cycle stays within the FP32 bar of the oracle (1e-3 relative L2; measured
    ~4e-6), and equals the library's FP32 batch run of the same cases
    (icempm_amd.batch) to the run-to-run reproducibility of the batched pass
    (hardware atomics; 1e-4); its snapshots hold the cycle's force bins within
    1e-5 of the oracle's (measured ~2e-7) and positions within 1e-9 m
    (measured ~1e-10)."""
    from icempm_amd._lib import BATCH_F32
    from icempm_amd.batch import batch_cases, build_batch
    print("cli: cm --batch, 3 cases, 1 cycle f32", flush=True)
    cases = batch_cases(3, 1500, 4000, seed=21)
    refs, N, hists = run_cm_batch(tmp_path, cases, True)
    runs = oracle_runs(refs, N)
    for k, (r, h) in enumerate(zip(runs, hists)):
        href = r[0]
        e = np.linalg.norm(h - href) / np.linalg.norm(href)
        print(f"case {k}: force rel L2 vs oracle {e:.2e}", flush=True)
        assert e <= 1e-3, (k, e)
    check_case_snapshots(tmp_path / "out", runs, N, 1e-5, 1e-9, tmp_path)
    b = build_batch(cases, BATCH_F32)
    b.run(N)
    assert b.error_flag() == 0
    for k, h in enumerate(hists):
        ref = b.force_history(k, N)
        e = np.linalg.norm(h - ref) / np.linalg.norm(ref)
        print(f"case {k}: force rel L2 vs library {e:.2e}", flush=True)
        assert e <= 1e-4, (k, e)
    b.close()


def test_cm_batch_ordered_is_bit_exact(gpu_available, tmp_path):
    """cm --batch with "precision": "ordered": the manifest's cases on the
    single-case ORDERED engine, per-grain strengths (GrainVariability 0.5,
    grain ids 0..999) on two of them; each case's snapshot (Points,
    Indenter_2D, step and time) and per-substep force equal the oracle run of
    that case alone bit for bit, as `cm --ordered` does."""
    import dataclasses

    from icempm_amd.batch import batch_cases
    print("cli: cm --batch ordered, 3 cases, 1 cycle", flush=True)
    cases = batch_cases(3, 1500, 4000, seed=23)
    cases = [dataclasses.replace(c, grain_variability=0.5) if k < 2 else c for k, c in enumerate(cases)]
    refs, N, hists = run_cm_batch(tmp_path, cases, False, precision="ordered", grain_seed=5)
    runs = oracle_runs(refs, N)
    crushed = []
    for k, ((href, bins, pts, p), h) in enumerate(zip(runs, hists)):
        assert np.array_equal(h, href), k
        snap = tmp_path / "out" / f"case_{k}" / "snapshots" / "d00001.h5"
        prm = SimParams.from_buffer_copy(binary(snap, "/Params", "u1", tmp_path).tobytes())
        assert prm.SimulationStep == N and prm.SimulationTime == p.SimulationTime
        assert np.array_equal(binary(snap, "/Indenter_2D", "<f8", tmp_path), bins), k
        dev = binary(snap, "/Points", "<f8", tmp_path).reshape(16, -1)
        assert np.array_equal(dev.view(np.uint64), pts.view(np.uint64)), k
        crushed.append(int(pts[0].view(np.uint8)[:p.nPts].sum()))
    print("crushed points per case", crushed, flush=True)
