"""Batches of independent cases (one launch per kernel per substep for all
cases, HIP graph replay) against the CPU oracle run case by case."""
import ctypes
import dataclasses

import numpy as np
import pytest

import oracle
from icempm_amd import IndentationCase, SimParams
from icempm_amd._lib import BATCH_F32, BATCH_F64, ICEMPM_E_ARG, ICEMPM_E_STATE, IcempmError
from icempm_amd.batch import Batch, batch_cases, build_batch, ingest_cases

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / n if n > 0 else np.linalg.norm(a - b)


def oracle_case(p0, soa0, nsub):
    p = SimParams()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(p0), ctypes.sizeof(SimParams))
    pts = soa0.copy()
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    hist, err = oracle.run(p, pts, grid, bins, nsub)
    assert err == 0
    return hist, pts


# Positions: FP64 storage keeps them to 1e-12 m.  FP32 storage holds them as
# 48-bit fixed point in units of h 2^-F, F = 48 - the bits of the larger grid
# dimension (quantum ~1e-14 m on these grids), and Fe - I in FP32 (~6e-8 of
# itself, which moves the stress by as much): 1e-10 m over 150 substeps
# (1e-8 of a cell).
F32_TOL_X = 1e-10


@pytest.mark.parametrize("precision,tol_f,tol_x", [(BATCH_F64, 1e-6, 1e-12), (BATCH_F32, 1e-3, F32_TOL_X)],
                         ids=["f64", "f32"])
def test_batch_matches_oracle_per_case(gpu_available, precision, tol_f, tol_x):
    print("batch", precision, flush=True)
    cases = batch_cases(3, 2000, 6000, seed=11)
    b = build_batch(cases, precision)
    nsub = 150
    b.run(nsub, use_graph=True)
    assert b.error_flag() == 0
    for c, case in enumerate(cases):
        p0, soa0 = b.initial[c]
        href, pref = oracle_case(p0, soa0, nsub)
        hist = b.force_history(c, nsub)
        n = case.n_points
        dev = b.points(c)
        assert np.abs(href).sum() > 0
        assert rel_l2(hist, href) <= tol_f, (c, rel_l2(hist, href))
        assert np.abs(dev[4:6, :n] - pref[4:6, :n]).max() <= tol_x
        assert np.array_equal(dev[0].view(np.uint8)[:n], pref[0].view(np.uint8)[:n])
    b.close()


def test_batch_graph_replay_matches_stream_run(gpu_available):
    print("batch graph", flush=True)
    cases = batch_cases(4, 1000, 4000, seed=12)
    g = build_batch(cases, BATCH_F64)
    s = build_batch(cases, BATCH_F64)
    for _ in range(3):                       # replays with host work between them
        g.run(40, use_graph=True)
        s.run(40, use_graph=False)
        for c in range(len(cases)):
            assert rel_l2(g.force_history(c, 40), s.force_history(c, 40)) <= 1e-9
    for c, case in enumerate(cases):
        n = case.n_points
        assert np.abs(g.points(c)[4:6, :n] - s.points(c)[4:6, :n]).max() <= 1e-12
    g.close()
    s.close()


@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "stream"])
def test_runs_continue_from_the_previous_runs_p2g(gpu_available, use_graph):
    """A run's last pass leaves the next substep's P2G in one of the two
    grids and the next run starts from it: runs of 7, 13, 1 and 9 substeps
    (odd lengths alternate the grid that holds it) follow the oracle's
    30-substep run as one run does, force history included; a restore, or a
    re-sort between runs, makes the next run start with a P2G of its own."""
    print("batch: runs continue from the previous P2G", use_graph, flush=True)
    cases = batch_cases(3, 1500, 4000, seed=14)
    b = build_batch(cases, BATCH_F64)
    lens = [7, 13, 1, 9]
    nsub = sum(lens)
    refs = [oracle_case(*b.initial[c], nsub) for c in range(len(cases))]
    b.checkpoint()

    def go(label, between=None):
        hist = [[] for _ in cases]
        for k, steps in enumerate(lens):
            if between is not None and k == 2:
                between()
            b.run(steps, use_graph=use_graph)
            for c in range(len(cases)):
                hist[c].append(b.force_history(c, steps))
        assert b.error_flag() == 0
        for c in range(len(cases)):
            assert rel_l2(np.concatenate(hist[c]), refs[c][0]) <= 1e-6, (label, c)
        check_state(label)

    def check_state(label):
        """positions to 1e-12 m; velocities, Fe and Bp (which a run's last
        pass stores for readback) to 1e-8 of each row's largest value"""
        for c, case in enumerate(cases):
            n = case.n_points
            pref = refs[c][1]
            dev = b.points(c)
            assert np.abs(dev[4:6, :n] - pref[4:6, :n]).max() <= 1e-12, (label, c)
            for r in (6, 7, 8, 9, 10, 11, 12, 13, 14, 15):
                scale = np.abs(pref[r, :n]).max()
                assert np.abs(dev[r, :n] - pref[r, :n]).max() <= 1e-8 * scale, (label, c, r)

    go("continued")
    b.restore()
    b.prepare(13)           # graphs of the 13-substep run captured ahead: the same results
    go("after a restore")
    b.restore()
    go("re-sorted between runs", between=b.resort)
    # a checkpoint taken after a run holds its velocities and Bp
    b.restore()
    b.run(10, use_graph=use_graph)
    b.checkpoint()
    b.run(20, use_graph=use_graph)
    h1 = [b.force_history(c, 20) for c in range(len(cases))]
    b.restore()
    b.run(20, use_graph=use_graph)
    for c in range(len(cases)):
        assert rel_l2(b.force_history(c, 20), refs[c][0][10:]) <= 1e-6, c
        assert rel_l2(h1[c], refs[c][0][10:]) <= 1e-6, c
    check_state("restored from a checkpoint taken after a run")
    b.close()


def test_resort_chunk_sizes_and_checkpoint_match_oracle(gpu_available):
    """Re-sorting between segments, other points-per-chunk values and a
    checkpoint/restore leave every case on the oracle's trajectory, and the
    points read back in their original order."""
    print("batch resort/chunk/checkpoint", flush=True)
    cases = batch_cases(3, 2000, 6000, seed=13)
    nsub = 120
    b = build_batch(cases, BATCH_F64)
    refs = [oracle_case(*b.initial[c], nsub) for c in range(len(cases))]

    def check(label):
        assert b.error_flag() == 0
        for c, case in enumerate(cases):
            n = case.n_points
            href, pref = refs[c]
            assert rel_l2(b.force_history(c, nsub), href) <= 1e-6, (label, c)
            assert np.abs(b.points(c)[4:6, :n] - pref[4:6, :n]).max() <= 1e-12, (label, c)

    b.checkpoint()
    b.set_tuning(chunk=64, resort_every=50)      # segments 50, 50, 20 with re-sorts between
    assert b.tuning() == (64, 50)
    b.run(nsub)
    check("chunk 64, re-sort every 50")
    b.restore()
    b.set_tuning(chunk=128, resort_every=0)
    b.resort()
    b.run(nsub, use_graph=False)
    check("chunk 128, re-sorted once")
    b.restore()
    rows = b.autotune(probe_steps=40)
    print("autotune", rows, "->", b.tuning(), flush=True)
    assert len(rows) == 9 and all(r[2] > 0 for r in rows)
    assert b.tuning() in {(r[0], r[1]) for r in rows}
    b.run(nsub)                                   # the state was restored by the tuner
    check("after autotune")
    b.close()


@pytest.mark.parametrize("precision", [BATCH_F64, BATCH_F32], ids=["f64", "f32"])
def test_batch_state_round_trip(gpu_available, precision):
    """A finalized batch holds every case's points in band order and, for FP32
    storage, positions in fixed point: read back before any substep, each case
    returns its ingested state in the original point order — exactly for FP64
    storage; for FP32 to within half a position quantum (gx h / 2^49) and the
    FP32 rounding of the other rows."""
    print("batch round trip", precision, flush=True)
    cases = batch_cases(4, 700, 5000, seed=17)
    b = build_batch(cases, precision)
    for c, case in enumerate(cases):
        p, soa0 = b.initial[c]
        n = case.n_points
        dev = b.points(c)
        assert np.array_equal(dev[0].view(np.uint8)[:n], soa0[0].view(np.uint8)[:n])   # crushed flags
        if precision == BATCH_F64:
            for r in range(3, 16):   # Jp_inv, position, velocity, Fe, Bp
                assert np.array_equal(dev[r, :n], soa0[r, :n]), r
        else:
            # half a quantum, the rounding of x / q (< 2^-5 of a quantum
            # for u < 2^48) and of the final scaling by q; the quantum is
            # h 2^-F with 48 - F the bits of the larger grid dimension
            ib = int(np.ceil(np.log2(max(p.GridX, p.GridY))))
            q = p.cellsize * 2.0 ** -(48 - ib)
            bound = (0.5 + 2.0 ** -5) * q + np.spacing(np.abs(soa0[4:6, :n]))
            assert (np.abs(dev[4:6, :n] - soa0[4:6, :n]) <= bound).all()
            for r in list(range(6, 16)) + [3]:
                ref = soa0[r, :n].astype(np.float32).astype(np.float64)
                if r in (8, 11, 3):   # Fe00, Fe11 and Jp_inv are stored less 1
                    ref = ((soa0[r, :n] - 1.0).astype(np.float32).astype(np.float64)) + 1.0
                assert np.array_equal(dev[r, :n], ref), r
    b.close()


# FP32 storage: the gathers read each node's velocity as an FP32 pair, a
# relative error of ~6e-8 that displaces a point by as much of its travel
# (0.03 m here), ~2e-9 m for most points.  The flight crushes ~90% of the
# points, and a point's FP32-stored plastic state can take it ~5e-7 m off
# the oracle's path (measured; which point depends on the atomic order).
# The skipped-block logic under test would be off by a cell if it failed.
@pytest.mark.parametrize("precision,tol_f,tol_x", [(BATCH_F64, 1e-6, 1e-12), (BATCH_F32, 1e-3, 1e-6)],
                         ids=["f64", "f32"])
def test_mass_entering_empty_grid_blocks(gpu_available, precision, tol_f, tol_x):
    """Every point starts with an upward velocity of 50 m/s, so over 300
    substeps the block rises by ~2.5 cells into grid rows that held no mass
    before: the grid update, which skips clearing blocks that took no mass,
    must still give the oracle's run."""
    print("batch: mass entering empty blocks", precision, flush=True)
    cases = batch_cases(2, 2000, 4000, seed=31)
    states = []
    for p, soa in ingest_cases(cases):
        soa = soa.copy()
        soa[7, :] = 50.0       # velocity y of every point (rows past the points are padding)
        states.append((p, soa))
    nsub = 300
    b = build_batch(cases, precision, states)
    b.run(nsub)
    assert b.error_flag() == 0
    for c, case in enumerate(cases):
        n = case.n_points
        href, pref = oracle_case(*states[c], nsub)
        dev = b.points(c)
        rise = pref[5, :n].mean() - states[c][1][5, :n].mean()
        assert rise > 2 * states[c][0].cellsize, rise
        e = rel_l2(b.force_history(c, nsub), href)
        dx = np.abs(dev[4:6, :n] - pref[4:6, :n]).max(axis=0)
        cr_d, cr_o = dev[0].view(np.uint8)[:n], pref[0].view(np.uint8)[:n]
        w = int(np.argmax(dx))
        print(f"case {c}: force {e:.2e}, max position difference {dx.max():.2e} m at point {w} "
              f"(crushed {cr_d[w]}/{cr_o[w]}), crushed flags differ at {np.flatnonzero(cr_d != cr_o)[:8]}, "
              f"oracle crushed {int(cr_o.sum())}", flush=True)
        assert e <= tol_f
        assert dx.max() <= tol_x
    b.close()


def test_grid_at_the_width_limit(gpu_available):
    """A case 32,767 cells wide, the widest grid a batch takes (cell indices
    travel as 16-bit halves in the particle pass; 2.1M points), follows the
    oracle for 20 substeps with FP64 and FP32 storage; a grid one cell wider
    is refused when the case is added."""
    case = IndentationCase(name="widest", ncx=32767 - 56, ncy=16, per_side=2, h=0.012, grid_x=32767,
                           grid_y=16 + 53, dt=2e-6, steps=20, ind_velocity=1.0)
    print("batch: widest grid", case.n_points, "points", flush=True)
    nsub = 20
    states = ingest_cases([case])
    href, pref = oracle_case(*states[0], nsub)
    n = case.n_points
    q = case.grid_x * case.h / 2.0 ** 48      # the FP32 position quantum on this grid, 1.4e-12 m
    for precision, tol_f, tol_x in ((BATCH_F64, 1e-6, 1e-12), (BATCH_F32, 1e-3, max(40 * q, F32_TOL_X))):
        b = build_batch([case], precision, states)
        b.run(nsub)
        assert b.error_flag() == 0
        e = rel_l2(b.force_history(0, nsub), href)
        dx = np.abs(b.points(0)[4:6, :n] - pref[4:6, :n]).max()
        print(f"precision {precision}: force rel L2 {e:.2e}, positions {dx:.2e} m", flush=True)
        assert np.abs(href).sum() > 0 and e <= tol_f
        assert dx <= tol_x
        b.close()
    small = batch_cases(1, 2000, 2000, seed=41)[0]
    too_wide = dataclasses.replace(small, grid_x=32768)
    with pytest.raises(IcempmError, match="32767"):
        build_batch([too_wide], BATCH_F64)


def test_batch_call_order_and_argument_errors(gpu_available):
    """The batch calls refuse work out of order and bad arguments with a
    status code and a message naming the problem, and the batch stays
    usable afterwards."""
    print("batch: call order and argument errors", flush=True)
    cases = batch_cases(2, 1500, 3000, seed=51)
    states = ingest_cases(cases)

    def refused(code, fragment, call, *args, **kw):
        with pytest.raises(IcempmError) as ei:
            call(*args, **kw)
        assert ei.value.code == code and fragment in str(ei.value), str(ei.value)

    b = Batch(BATCH_F64)
    refused(ICEMPM_E_ARG, "", b.finalize)                        # no case yet
    refused(ICEMPM_E_STATE, "not finalized", b.run, 10)
    refused(ICEMPM_E_STATE, "not finalized", b.checkpoint)
    p, soa = states[0]
    other_dt = states[1][0].copy()
    other_dt.InitialTimeStep = p.InitialTimeStep * 2
    b.add_state(cases[0], p, soa)
    refused(ICEMPM_E_ARG, "time step", b.add_state, cases[1], other_dt, states[1][1])
    # sizes and derived constants the kernels index and scale by are checked
    # on the host, before any kernel runs
    bad = states[1][0].copy()
    bad.GridTotal = bad.GridX * bad.GridY - 1
    refused(ICEMPM_E_ARG, "GridTotal", b.add_state, cases[1], bad, states[1][1])
    bad = states[1][0].copy()
    bad.cellsize_inv = 1.01 / bad.cellsize
    refused(ICEMPM_E_ARG, "consistent", b.add_state, cases[1], bad, states[1][1])
    bad = states[1][0].copy()
    bad.cellsize = -bad.cellsize
    refused(ICEMPM_E_ARG, "positive", b.add_state, cases[1], bad, states[1][1])
    bad = states[1][0].copy()
    bad.n_indenter_subdivisions = 0
    refused(ICEMPM_E_ARG, "n_indenter_subdivisions", b.add_state, cases[1], bad, states[1][1])
    b.add_state(cases[1], *states[1])
    b.finalize()
    refused(ICEMPM_E_STATE, "already finalized", b.finalize)
    refused(ICEMPM_E_STATE, "already finalized", b.add_state, cases[0], p, soa)
    refused(ICEMPM_E_ARG, "64, 128 or 256", b.set_tuning, chunk=100)
    refused(ICEMPM_E_ARG, ">= 0", b.set_tuning, resort_every=-1)
    refused(ICEMPM_E_ARG, "positive capacity", b.set_harvest, 0)
    refused(ICEMPM_E_ARG, "", b.force_history, 2, 1)            # no case 2
    b.run(30)
    refused(ICEMPM_E_ARG, "", b.force_history, 0, 31)           # longer than the last run
    # still usable: the run equals the oracle
    href, pref = oracle_case(*states[0], 30)
    assert rel_l2(b.force_history(0, 30), href) <= 1e-6
    assert b.error_flag() == 0
    b.close()


@pytest.mark.parametrize("precision,tol_f,tol_x", [(BATCH_F64, 1e-6, 1e-12), (BATCH_F32, 1e-3, 2e-8)],
                         ids=["f64", "f32"])
def test_tiny_and_ragged_cases(gpu_available, precision, tol_f, tol_x):
    """Cases of 4, 60, 272 and 264 points in one batch (one cell; a case
    smaller than a wavefront; cases that end one cell row or a few points
    into a chunk) each follow the oracle."""
    print("batch: tiny and ragged cases", precision, flush=True)
    shapes = [(1, 1), (3, 5), (17, 4), (33, 2)]
    cases = [IndentationCase(name=f"ragged{i}", ncx=nx, ncy=ny, per_side=2, h=0.012, grid_x=nx + 56,
                             grid_y=ny + 53, dt=2e-6, steps=50, ind_diameter=0.05, ind_velocity=0.5,
                             ind_depth=0.01, seed=0x1CE5EED + i)
             for i, (nx, ny) in enumerate(shapes)]
    assert [c.n_points for c in cases] == [4, 60, 272, 264]
    nsub = 50
    b = build_batch(cases, precision)
    b.run(nsub)
    assert b.error_flag() == 0
    for c, case in enumerate(cases):
        n = case.n_points
        href, pref = oracle_case(*b.initial[c], nsub)
        assert rel_l2(b.force_history(c, nsub), href) <= tol_f, c
        assert np.abs(b.points(c)[4:6, :n] - pref[4:6, :n]).max() <= tol_x, c
    b.close()


def grain_states(cases, gv=0.5, seed=5):
    """Ingested states with per-point grain ids drawn from [-40, 1000) and
    GrainVariability gv (the reference's default material); the last case
    keeps a uniform material, so one batch mixes both."""
    rng = np.random.default_rng(seed)
    out = []
    for i, (p, soa) in enumerate(ingest_cases(cases)):
        p, soa = p.copy(), soa.copy()
        if i < len(cases) - 1:
            n = p.nPts
            pitch = soa.shape[1]
            soa[0].view(np.uint8)[pitch:pitch + 2 * n].view(np.int16)[:] = rng.integers(-40, 1000, size=n)
            p.GrainVariability = gv
        out.append((p, soa))
    return out


@pytest.mark.parametrize("precision,tol_f,tol_x", [(BATCH_F64, 1e-6, 1e-12), (BATCH_F32, 1e-3, F32_TOL_X)],
                         ids=["f64", "f32"])
def test_batch_grain_variability_matches_oracle(gpu_available, precision, tol_f, tol_x):
    """Cases with per-grain strengths (GetParametersForGrain,
    gpu_implementation5.cu:97-111) next to a uniform one: each follows the
    oracle's run of it, crushed flags included, and the grain ids read back
    unchanged -- also after a re-sort and a checkpoint/restore, which move
    the grain row with the points."""
    print("batch: grain variability", precision, flush=True)
    cases = batch_cases(3, 2000, 6000, seed=61)
    states = grain_states(cases)
    nsub = 200
    refs = [oracle_case(*states[c], nsub) for c in range(len(cases))]
    crushed = [int(r[1][0].view(np.uint8)[:cases[c].n_points].sum()) for c, r in enumerate(refs)]
    print("oracle crushed points per case", crushed, flush=True)
    assert min(crushed[:-1]) >= 5
    b = build_batch(cases, precision, states)

    def check(label):
        assert b.error_flag() == 0
        for c, case in enumerate(cases):
            n = case.n_points
            href, pref = refs[c]
            dev = b.points(c)
            e = rel_l2(b.force_history(c, nsub), href)
            assert e <= tol_f, (label, c, e)
            assert np.abs(dev[4:6, :n] - pref[4:6, :n]).max() <= tol_x, (label, c)
            assert np.array_equal(dev[0].view(np.uint8)[:n], pref[0].view(np.uint8)[:n]), (label, c)
            pitch = dev.shape[1]
            g_dev = dev[0].view(np.uint8)[pitch:pitch + 2 * n].view(np.int16)
            s0 = states[c][1]
            g_ref = s0[0].view(np.uint8)[s0.shape[1]:s0.shape[1] + 2 * n].view(np.int16)
            assert np.array_equal(g_dev, g_ref), (label, c)

    b.checkpoint()
    b.run(nsub)
    check("straight")
    b.restore()
    b.set_tuning(resort_every=70)
    b.run(nsub)
    check("re-sorted every 70")
    b.close()


@pytest.mark.parametrize("precision,tol", [(BATCH_F64, 1e-6), (BATCH_F32, 1e-3)], ids=["f64", "f32"])
def test_batch_force_bins_match_oracle(gpu_available, precision, tol):
    """Each case's angular indenter-force bins (Indenter_2D,
    gpu_implementation5.cu:262-269), summed over a run, against the oracle's
    bins over the same substeps; they also sum to the run's force history."""
    print("batch: force bins", precision, flush=True)
    cases = batch_cases(3, 2000, 6000, seed=71)
    nsub = 300
    b = build_batch(cases, precision)
    b.run(nsub)
    assert b.error_flag() == 0
    for c, case in enumerate(cases):
        p0, soa0 = b.initial[c]
        p = SimParams()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(p0), ctypes.sizeof(SimParams))
        assert nsub < p.UpdateEveryNthStep      # one cycle: the oracle's bins sum over all of it
        pts = soa0.copy()
        grid = oracle.new_grid(p)
        bins = np.zeros(2 * p.n_indenter_subdivisions)
        oracle.run(p, pts, grid, bins, nsub)
        got = b.force_bins(c)
        assert np.count_nonzero(bins) >= 4           # several bins in contact
        assert np.array_equal(got != 0, bins != 0) or precision == BATCH_F32
        e = rel_l2(got, bins)
        assert e <= tol, (c, e)
        tot = b.force_history(c, nsub).sum(axis=0)
        assert np.allclose(got.reshape(-1, 2).sum(axis=0), tot, rtol=1e-9, atol=1e-9 * np.abs(tot).max())
    # a second run starts its bins from zero
    b.run(10)
    for c in range(len(cases)):
        tot = b.force_history(c, 10).sum(axis=0)
        assert np.allclose(b.force_bins(c).reshape(-1, 2).sum(axis=0), tot, rtol=1e-9,
                           atol=1e-9 * max(np.abs(tot).max(), 1e-30))
    b.close()
