"""Pins the host-loop, derived-constant, cell-key and ingest parts of the
oracle (model.cpp:11-47, parameters_sim.cpp:121-147,
snapshotmanager.cpp:186-304)."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from icempm_amd import INDENT_TINY, SimParams, configure_params, host, lattice_points


@pytest.mark.parametrize("dt,n", [(1e-5, 500), (5e-6, 1000), (2e-6, 2500), (1e-6, 5000), (3e-5, 166)])
def test_update_every_nth_step(dt, n):
    p = SimParams()
    host().icempm_simparams_reset(ctypes.byref(p))
    p.InitialTimeStep = dt
    host().icempm_simparams_derive(ctypes.byref(p))
    assert p.UpdateEveryNthStep == n


def test_lame_constants():
    p = SimParams()
    host().icempm_simparams_reset(ctypes.byref(p))
    p.YoungsModulus, p.PoissonsRatio = 9e9, 0.3
    host().icempm_simparams_derive(ctypes.byref(p))
    assert p.mu == pytest.approx(3.4615384615e9, rel=1e-10)
    assert p.lambda_ == pytest.approx(5.1923076923e9, rel=1e-10)
    assert p.kappa == pytest.approx(7.5e9, rel=1e-14)


def test_cell_keys_on_cell_centres():
    gx, h = 37, 0.25
    cx, cy = np.meshgrid(np.arange(30), np.arange(20), indexing="xy")
    # points a quarter cell off each node -> the key is the node index
    xy = np.stack([(cx.ravel() + 0.75) * h, (cy.ravel() + 0.75) * h], axis=1).astype(np.float32)
    keys = oracle.cell_keys_f32(xy, 1 / h, gx)
    assert np.array_equal(keys, cx.ravel() + cy.ravel() * gx)


def test_cell_key_is_not_contracted():
    # Near cell boundaries, trunc(round(x*hinv) - 0.5) and the fused
    # trunc(round(x*hinv - 0.5)) can differ.  The key is defined with two
    # roundings (no FMA).  Double positions near (k + 0.5) h give such
    # inputs; check the oracle takes the two-rounding form.
    from fractions import Fraction
    hinv = 1 / 0.012
    xs, twos = [], []
    for k in range(1, 2000):
        x0 = (k + 0.5) / hinv
        x = x0
        for _ in range(8):
            x = np.nextafter(x, -np.inf)
        for _ in range(17):
            xd = float(x)
            two = int(xd * hinv - 0.5)
            fused = int(float(Fraction(xd) * Fraction(hinv) - Fraction(1, 2)))
            if fused != two:
                xs.append(xd)
                twos.append(two)
            x = np.nextafter(x, np.inf)
    assert len(xs) >= 2          # e.g. x = 0.054 and x = 3.078
    pts = np.zeros((16, len(xs) + 64))
    pts[4, :len(xs)] = xs
    keys = oracle.cell_keys_soa(pts, len(xs), hinv, 10)
    assert keys.tolist() == twos


def test_stable_order_matches_numpy():
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 50, size=5000).astype(np.int32)
    assert np.array_equal(oracle.stable_order(keys), np.argsort(keys, kind="stable"))


def test_run_time_accumulation_and_cycle_bins():
    case = INDENT_TINY
    p = SimParams()
    configure_params(p, case)
    p.InitialTimeStep = 1e-4          # UpdateEveryNthStep = 50
    host().icempm_simparams_derive(ctypes.byref(p))
    xy, g = lattice_points(case)
    pts = oracle.ingest(p, xy, g, case.volume, pitch=((case.n_points + 63) // 64) * 64)
    p.nGridPitch = ((p.GridTotal + 63) // 64) * 64
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    N = p.UpdateEveryNthStep
    assert N == 50
    hist, err = oracle.run(p, pts, grid, bins, 70)
    assert err == 0
    t = 0.0
    for _ in range(70):
        t += p.InitialTimeStep
    assert p.SimulationTime == t           # repeated addition, bit for bit
    assert p.SimulationStep == 70
    assert p.indenter_x == p.indenter_x_initial + t * p.IndVelocity
    # bins hold the sum over the current cycle only (substeps 50..69)
    assert np.allclose(bins.reshape(-1, 2).sum(axis=0), hist[50:].sum(axis=0), rtol=1e-9, atol=1e-9)


def test_ingest_places_block_and_indenter():
    case = INDENT_TINY
    p = SimParams()
    configure_params(p, case)
    xy, g = lattice_points(case)
    pitch = ((case.n_points + 63) // 64) * 64
    pts = oracle.ingest(p, xy, g, case.volume, pitch)
    n = case.n_points
    h = p.cellsize
    # sorted by the raw cell key
    keys = oracle.cell_keys_f32(np.stack([pts[4, :n], pts[5, :n]], 1).astype(np.float32), 1 / h, p.GridX)
    assert (np.diff(oracle.cell_keys_f32(xy[oracle.stable_order(oracle.cell_keys_f32(xy, p.cellsize_inv, p.GridX))], p.cellsize_inv, p.GridX)) >= 0).all()
    # block centred in x, lifted by 2h
    x_offset = (p.GridX * h - (p.xmax - p.xmin)) / 2
    assert pts[4, :n].min() == pytest.approx(p.xmin + x_offset, abs=1e-6)
    assert pts[4, :n].max() == pytest.approx(p.xmax + x_offset, abs=1e-6)
    assert pts[5, :n].min() == pytest.approx(p.ymin + 2 * h, abs=1e-6)
    # disc touches the block's top-left corner at depth IndDepth
    r = p.IndDiameter / 2
    assert p.indenter_y == pytest.approx(p.ymax + 2 * h + r - p.IndDepth, rel=1e-15)
    assert p.indenter_x % h == pytest.approx(0, abs=1e-12) or abs(p.indenter_x % h - h) < 1e-12
    assert p.ParticleMass == pytest.approx(case.volume / n * p.Density, rel=1e-7)
    assert (pts[8, :n] == 1).all() and (pts[11, :n] == 1).all() and (pts[3, :n] == 1).all()
    assert keys.shape == (n,)


def test_timing_build_equals_parity_build():
    """The CPU baseline's -O3/AVX2 build of the oracle gives the parity
    build's results bit for bit (both keep a*b+c uncontracted, and neither
    reassociates)."""
    import ctypes
    from icempm_amd import INDENT_TINY, SimParams, configure_params, lattice_points
    p = SimParams()
    configure_params(p, INDENT_TINY)
    xy, g = lattice_points(INDENT_TINY)
    pts = oracle.ingest(p, xy, g, INDENT_TINY.volume, pitch=2432)
    p.nGridPitch = (p.GridTotal + 63) // 64 * 64
    q = SimParams()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(SimParams))
    pts2 = pts.copy()
    h1, e1 = oracle.run(p, pts, oracle.new_grid(p), np.zeros(200), 300)
    h2, e2 = oracle.run(q, pts2, oracle.new_grid(q), np.zeros(200), 300, nthreads=1, fast=True)
    assert e1 == 0 and e2 == 0
    assert np.array_equal(h1, h2)
    assert np.array_equal(pts.view(np.uint64), pts2.view(np.uint64))
