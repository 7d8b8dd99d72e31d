"""HIP path vs the CPU oracle on the same seeded inputs.

ORDERED mode must reproduce the oracle bit for bit (same operation order,
correctly rounded FP64).  FAST mode differs only in the order of the P2G and
indenter-force sums; it is held to the FP64 tolerance (indenter force
history relative L2 <= 1e-6 in FP64) over the pre-fracture horizon and to
tight per-field bounds.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from icempm_amd import (INDENT_1M6, INDENT_60K, INDENT_TINY, MODE_FAST, MODE_ORDERED, Model, SimParams,
                        configure_params, lattice_points)
from icempm_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def load_pair(case, mode, dt=None):
    """The product's ingest on the GPU and the oracle's on the CPU."""
    m = Model(mode)
    configure_params(m.prms, case)
    if dt is not None:
        m.prms.InitialTimeStep = dt
        _lib.host().icempm_simparams_derive(ctypes.byref(m.prms))
    xy, g = lattice_points(case)
    m.load_raw_points(xy, g, case.volume)
    p = SimParams()
    configure_params(p, case)
    if dt is not None:
        p.InitialTimeStep = dt
        _lib.host().icempm_simparams_derive(ctypes.byref(p))
    pts = oracle.ingest(p, xy, g, case.volume, pitch=m.prms.nPtsPitch)
    p.nGridPitch = m.prms.nGridPitch
    return m, p, pts


def rel_l2(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    n = np.linalg.norm(np.asarray(b))
    return d / n if n > 0 else d


@pytest.mark.parametrize("case", [INDENT_TINY, INDENT_60K], ids=lambda c: c.name)
def test_ingest_bit_exact(gpu_available, case):
    say("ingest", case.name)
    m, p, pts = load_pair(case, MODE_FAST)
    dev = m.points()
    assert np.array_equal(dev.view(np.uint64), pts.view(np.uint64))
    q = m.prms
    for f in ("nPts", "Volume", "xmin", "xmax", "ymin", "ymax", "nxmin", "nxmax", "nymin",
              "nymax", "indenter_x", "indenter_y", "indenter_x_initial", "indenter_y_initial",
              "ParticleVolume", "ParticleMass"):
        assert getattr(q, f) == getattr(p, f), f
    m.close()


def test_device_hello(gpu_available):
    say("hello kernel")
    m = Model()
    _lib.check_ctx(_lib.hip().icempm_test(m.ctx), m.ctx)
    m.close()


@pytest.mark.parametrize("case,nsub", [(INDENT_TINY, 400), (INDENT_60K, 120)],
                         ids=lambda x: getattr(x, "name", str(x)))
def test_ordered_run_bit_exact(gpu_available, case, nsub):
    say("ordered run", case.name, nsub)
    m, p, pts = load_pair(case, MODE_ORDERED)
    hist_dev = m.run(nsub)
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    hist_ref, err = oracle.run(p, pts, grid, bins, nsub)
    assert err == 0 and m.error_flag() == 0
    dev = m.points()
    n = case.n_points
    bad = np.flatnonzero((dev[:, :n] != pts[:, :n]).any(axis=0))
    assert bad.size == 0, f"{bad.size} particles differ, first {bad[:5]}"
    assert np.array_equal(dev.view(np.uint64), pts.view(np.uint64))
    assert np.array_equal(hist_dev, hist_ref)
    assert np.array_equal(m.bins(), bins)
    assert np.abs(hist_ref).sum() > 0          # the disc is in contact
    assert m.prms.SimulationTime == p.SimulationTime
    m.close()


def test_ordered_single_operations_bit_exact(gpu_available):
    """The reference's four per-substep calls one by one (cuda_reset_grid,
    cuda_p2g, cuda_update_nodes, cuda_g2p) after a warm-up."""
    say("ordered single ops")
    case = INDENT_TINY
    m, p, pts = load_pair(case, MODE_ORDERED)
    m.run(50)
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    oracle.run(p, pts, grid, bins, 50)
    H = _lib.hip()
    ctx = m.ctx
    _lib.check_ctx(H.icempm_reset_grid(ctx), ctx)
    _lib.check_ctx(H.icempm_p2g(ctx), ctx)
    g_ref = oracle.new_grid(p)
    oracle.p2g(p, pts, g_ref)
    assert np.array_equal(m.grid(), g_ref)
    ix, iy = p.indenter_x + 0.01, p.indenter_y
    _lib.check_ctx(H.icempm_update_nodes(ctx, ix, iy), ctx)
    oracle.update_nodes(p, g_ref, bins, ix, iy)
    assert np.array_equal(m.grid(), g_ref)
    assert np.array_equal(m.bins(), bins)
    _lib.check_ctx(H.icempm_g2p(ctx, 1), ctx)
    oracle.g2p(p, pts, g_ref, True)
    assert np.array_equal(m.points(), pts)
    m.close()


@pytest.mark.parametrize("case,nsub", [(INDENT_TINY, 400), (INDENT_60K, 150)],
                         ids=lambda x: getattr(x, "name", str(x)))
def test_fast_within_tolerance(gpu_available, case, nsub):
    say("fast run", case.name, nsub)
    m, p, pts = load_pair(case, MODE_FAST)
    hist_dev = m.run(nsub)
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    hist_ref, err = oracle.run(p, pts, grid, bins, nsub)
    assert err == 0 and m.error_flag() == 0
    n = case.n_points
    dev = m.points()
    # the FP64 tolerance: indenter force history within 1e-6 rel L2
    assert rel_l2(hist_dev, hist_ref) <= 1e-6
    assert rel_l2(m.bins(), bins) <= 1e-6
    # positions to ~1e-12 of the domain, deformation gradient to 1e-9
    assert np.abs(dev[4:6, :n] - pts[4:6, :n]).max() <= 1e-12
    assert np.abs(dev[8:12, :n] - pts[8:12, :n]).max() <= 1e-9
    assert np.abs(dev[3, :n] - pts[3, :n]).max() <= 1e-9
    # the plastic state (crushed flags) is identical over this horizon
    assert np.array_equal(dev[0].view(np.uint8)[:n], pts[0].view(np.uint8)[:n])
    m.close()


def test_fast_p2g_single_call(gpu_available):
    say("fast p2g")
    case = INDENT_60K
    m, p, pts = load_pair(case, MODE_FAST)
    grid = oracle.new_grid(p)
    oracle.run(p, pts, grid, np.zeros(200), 30)
    # same state on both sides: upload the oracle's points
    m.host_points()[:] = pts
    H = _lib.hip()
    _lib.check_ctx(H.icempm_transfer_points_to_device(m.ctx), m.ctx)
    _lib.check_ctx(H.icempm_reset_grid(m.ctx), m.ctx)
    _lib.check_ctx(H.icempm_p2g(m.ctx), m.ctx)
    g_ref = oracle.new_grid(p)
    oracle.p2g(p, pts, g_ref)
    g = m.grid()
    # FAST mode sums in another order and contracts a*b+c into FMAs; the
    # stress term kappa*(J*J-1) with J ~ 1-1e-5 loses ~5 digits without FMA,
    # so per-node differences reach ~1e-11 of the row scale.  A dropped or
    # doubled contribution would be ~1e-3.
    scale = np.abs(g_ref).max(axis=1)
    err = np.abs(g - g_ref).max(axis=1) / scale
    say("fast p2g: max error / row scale", err)
    assert (err <= 1e-9).all()
    assert np.array_equal(g[0] != 0, g_ref[0] != 0)
    assert g[0].sum() == pytest.approx(g_ref[0].sum(), rel=1e-13)
    m.close()


def test_model_step_cycle_ordered(gpu_available):
    """Model::Step (one output cycle of UpdateEveryNthStep substeps, then the
    async readback) against the oracle, bit for bit."""
    say("model step")
    case = INDENT_TINY
    m, p, pts = load_pair(case, MODE_ORDERED)
    N = m.prms.UpdateEveryNthStep
    assert N == 2500
    cont = m.step()
    m.wait_cycle()
    assert cont
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    oracle.run(p, pts, grid, bins, N)
    hp = m.host_points()
    assert np.array_equal(hp.view(np.uint64), pts.view(np.uint64))
    assert np.array_equal(m.host_force(), bins)
    assert m.prms.SimulationStep == N and m.prms.SimulationTime == p.SimulationTime
    assert m.prms.indenter_x == p.indenter_x
    # P, Q were recorded on the cycle's last substep
    assert np.abs(hp[2, :case.n_points]).max() > 0
    m.close()


def test_out_of_grid_flag_stops_the_run(gpu_available):
    say("out of grid")
    case = INDENT_TINY
    m, p, pts = load_pair(case, MODE_FAST)
    buf = m.points()
    buf[4, 0] = -0.05                        # no base cell
    hp = m.host_points()
    hp[:] = buf
    _lib.check_ctx(_lib.hip().icempm_transfer_points_to_device(m.ctx), m.ctx)
    m.run(2)
    assert m.error_flag() == 1
    # Model::Step reports it through error_code on the same cycle: the
    # readback's flag copy has landed when cuda_transfer_from_device returns
    cont = m.step()
    m.wait_cycle()
    assert not cont
    assert np.isfinite(m.points()[4:6, 1:case.n_points]).all()
    m.close()


def test_call_order_errors(gpu_available):
    say("call order")
    H = _lib.hip()
    ctx = ctypes.c_void_p()
    assert H.icempm_create(ctypes.byref(ctx)) == 0
    assert H.icempm_p2g(ctx) == _lib.ICEMPM_E_STATE
    assert b"constants" in H.icempm_last_error(ctx)
    assert H.icempm_update_constants(ctx, ctypes.byref(SimParams()), 100) == _lib.ICEMPM_E_ARG
    H.icempm_destroy(ctx)


def test_full_size_case(gpu_available):
    """The 1.6M-point configuration itself: ORDERED bit-exact for a few
    substeps, FAST within the FP64 tolerance, and the size-independent
    conservation of a P2G (total mass N m; total momentum m sum v, since the
    APIC affine term sums to zero over the stencil)."""
    case = INDENT_1M6
    say("full size ordered", case.name)
    m, p, pts = load_pair(case, MODE_ORDERED)
    m.run(6)
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    _, err = oracle.run(p, pts, grid, bins, 6)
    assert err == 0 and m.error_flag() == 0
    assert np.array_equal(m.points().view(np.uint64), pts.view(np.uint64))
    m.close()

    say("full size fast", case.name)
    m, p, pts = load_pair(case, MODE_FAST)
    nsub = 30
    hist_dev = m.run(nsub)
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    hist_ref, err = oracle.run(p, pts, grid, bins, nsub)
    assert err == 0 and m.error_flag() == 0
    n = case.n_points
    dev = m.points()
    assert rel_l2(hist_dev, hist_ref) <= 1e-6
    assert np.abs(dev[4:6, :n] - pts[4:6, :n]).max() <= 1e-12
    assert np.abs(dev[8:12, :n] - pts[8:12, :n]).max() <= 1e-9
    assert np.array_equal(dev[0].view(np.uint8)[:n], pts[0].view(np.uint8)[:n])

    say("full size p2g conservation", case.name)
    H = _lib.hip()
    _lib.check_ctx(H.icempm_reset_grid(m.ctx), m.ctx)
    _lib.check_ctx(H.icempm_p2g(m.ctx), m.ctx)
    g = m.grid()
    mass = m.prms.ParticleMass
    assert g[0].sum() == pytest.approx(n * mass, rel=1e-12)
    mom = mass * dev[6:8, :n].sum(axis=1)
    scale = mass * np.abs(dev[6:8, :n]).sum() + 1e-300
    assert np.abs(g[1:3].sum(axis=1) - mom).max() <= 1e-9 * scale
    m.close()
