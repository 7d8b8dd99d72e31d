"""Pins the oracle's P2G, grid update and G2P (gpu_implementation5.cu:148-379)
with the B-spline / APIC identities and per-node closed forms, using an
independent textbook form of the quadratic B-spline."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from icempm_amd import SimParams, host, nacc_strengths


def make_params(gx=16, gy=16, h=0.1, dt=1e-4, **kw):
    p = SimParams()
    host().icempm_simparams_reset(ctypes.byref(p))
    C, T, S = nacc_strengths(1.2, 0.8, 1e6)
    p.IceCompressiveStrength, p.IceTensileStrength, p.IceShearStrength = C, T, S
    p.YoungsModulus = 9e9
    p.GrainVariability = 0.0
    p.GridX, p.GridY = gx, gy
    p.GridXDimension = gx * h
    p.InitialTimeStep = dt
    p.IndDiameter = 0.3
    for k, v in kw.items():
        setattr(p, k, v)
    host().icempm_simparams_derive(ctypes.byref(p))
    p.nGridPitch = ((p.GridTotal + 63) // 64) * 64
    return p


def make_points(p, pos, vel=None, Fe=None, Bp=None, vol=1e-3):
    n = len(pos)
    pitch = max(64, ((n + 63) // 64) * 64)
    p.nPts = n
    p.nPtsPitch = pitch
    p.ParticleVolume = vol
    p.ParticleMass = vol * p.Density
    pts = np.zeros((16, pitch))
    pts[3, :n] = 1
    pts[4:6, :n] = np.asarray(pos, dtype=np.float64).T
    if vel is not None:
        pts[6:8, :n] = np.asarray(vel, dtype=np.float64).T
    F = np.tile(np.eye(2).reshape(4, 1), (1, n)) if Fe is None else np.asarray(Fe).reshape(n, 4).T
    pts[8:12, :n] = F
    if Bp is not None:
        pts[12:16, :n] = np.asarray(Bp).reshape(n, 4).T
    return pts


def bspline(r):
    """Quadratic B-spline N(r), r in cell units (textbook piecewise form)."""
    r = abs(r)
    if r < 0.5:
        return 0.75 - r * r
    if r < 1.5:
        return 0.5 * (1.5 - r) ** 2
    return 0.0


def stencil_nodes(p, x):
    """All (i, j, w) with nonzero weight, from the piecewise B-spline."""
    h = p.cellsize
    out = []
    for i in range(p.GridX):
        wx = bspline(x[0] / h - i)
        if wx == 0:
            continue
        for j in range(p.GridY):
            wy = bspline(x[1] / h - j)
            if wy:
                out.append((i, j, wx * wy))
    return out


# ---- P2G -------------------------------------------------------------------------
def test_single_particle_brute_force():
    p = make_params()
    rng = np.random.default_rng(5)
    for _ in range(20):
        x = rng.uniform(0.4, 1.2, size=2)
        v = rng.normal(size=2)
        Bp = rng.normal(size=(2, 2))
        pts = make_points(p, [x], [v], Bp=[Bp])
        grid = oracle.new_grid(p)
        assert oracle.p2g(p, pts, grid) == 0
        tau = oracle.kirchhoff(np.eye(2), p.mu, p.kappa)
        A = p.ParticleMass * Bp - p.InitialTimeStep * p.ParticleVolume * p.Dp_inv * tau
        m = p.ParticleMass
        expect = np.zeros_like(grid)
        nodes = stencil_nodes(p, x)
        assert len(nodes) == 9
        for i, j, w in nodes:
            k = i + j * p.GridX
            d = np.array([i * p.cellsize, j * p.cellsize]) - x
            expect[0, k] += w * m
            expect[1:, k] += w * (m * v + A @ d)
        assert np.allclose(grid, expect, rtol=1e-12, atol=1e-15)


def test_p2g_moments():
    # sum w = 1, sum w (x_i - x_p) = 0, sum w (x_i - x_p)(x_i - x_p)^T = h^2/4 I:
    # mass, momentum and the affine moment of APIC are transferred exactly
    p = make_params()
    rng = np.random.default_rng(6)
    x = rng.uniform(0.4, 1.2, size=(50, 2))
    v = rng.normal(size=(50, 2))
    Bp = rng.normal(size=(50, 2, 2))
    F = np.eye(2) + 1e-3 * rng.normal(size=(50, 2, 2))
    pts = make_points(p, x, v, Fe=F, Bp=Bp)
    grid = oracle.new_grid(p)
    assert oracle.p2g(p, pts, grid) == 0
    m = p.ParticleMass
    assert grid[0].sum() == pytest.approx(50 * m, rel=1e-13)
    assert np.allclose(grid[1:].sum(axis=1), m * v.sum(axis=0), rtol=1e-12, atol=1e-12)
    # single particle: sum_i p_i (x_i - x_p)^T = A h^2/4
    pts1 = make_points(p, x[:1], v[:1], Fe=F[:1], Bp=Bp[:1])
    g1 = oracle.new_grid(p)
    oracle.p2g(p, pts1, g1)
    tau = oracle.kirchhoff(F[0], p.mu, p.kappa)
    A = m * Bp[0] - p.InitialTimeStep * p.ParticleVolume * p.Dp_inv * tau
    M = np.zeros((2, 2))
    for i, j, _ in stencil_nodes(p, x[0]):
        k = i + j * p.GridX
        d = np.array([i, j]) * p.cellsize - x[0]
        M += np.outer(g1[1:, k], d)
    assert np.allclose(M, A * p.cellsize ** 2 / 4, rtol=1e-10, atol=1e-14)


def test_p2g_out_of_grid_sets_flag_and_skips():
    p = make_params()
    # base index = (int)(x/h - 0.5) truncates toward zero: x = 0.02 still
    # maps to base 0 (inside); x = 1.55 maps to base 15 whose stencil
    # reaches nodes 16 and 17 (outside a 16-wide grid); x = -0.1 has no base
    assert oracle.p2g(p, make_points(p, [[0.02, 0.5]]), oracle.new_grid(p)) == 0
    for x in ([1.55, 0.5], [-0.1, 0.5]):
        grid = oracle.new_grid(p)
        assert oracle.p2g(p, make_points(p, [x]), grid) == 1
        assert np.isfinite(grid).all()
        assert grid[0].sum() < p.ParticleMass      # the outside part is dropped


# ---- grid update --------------------------------------------------------------------
def node_grid(p, k, m, mom):
    g = oracle.new_grid(p)
    g[0, k] = m
    g[1:, k] = mom
    return g


def test_update_free_node_gravity():
    p = make_params(Gravity=9.81)
    k = 7 + 8 * p.GridX
    g = node_grid(p, k, 2.0, [0.4, 0.2])
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    oracle.update_nodes(p, g, bins, -10.0, -10.0)
    assert g[1, k] == 0.2
    assert g[2, k] == pytest.approx(0.1 - p.InitialTimeStep * 9.81, rel=1e-15)
    assert not bins.any()


def test_update_clamps_speed():
    p = make_params(Gravity=0.0)
    vmax = 0.5 * p.cellsize / p.InitialTimeStep
    k = 7 + 8 * p.GridX
    g = node_grid(p, k, 1.0, [3 * vmax, 4 * vmax])
    oracle.update_nodes(p, g, np.zeros(200), -10.0, -10.0)
    assert np.hypot(g[1, k], g[2, k]) == pytest.approx(vmax, rel=1e-14)
    assert g[1, k] / g[2, k] == pytest.approx(0.75, rel=1e-14)


@pytest.mark.parametrize("ix,iy,v,expect", [
    (5, 0, [1.0, 1.0], [0.0, 0.0]),           # bottom rows fixed
    (5, 2, [1.0, -1.0], [0.0, 0.0]),
    (5, 13, [0.5, 1.0], [0.5, 0.0]),           # top: no upward motion
    (5, 13, [0.5, -1.0], [0.5, -1.0]),
    (1, 8, [-1.0, 0.3], [0.0, 0.3]),           # left wall: no leftward motion
    (1, 8, [1.0, 0.3], [1.0, 0.3]),
    (14, 8, [1.0, 0.3], [0.0, 0.3]),           # right wall
    (14, 8, [-1.0, 0.3], [-1.0, 0.3]),
])
def test_update_boundary_regions(ix, iy, v, expect):
    p = make_params(Gravity=0.0)
    k = ix + iy * p.GridX
    g = node_grid(p, k, 1.0, v)
    oracle.update_nodes(p, g, np.zeros(200), -10.0, -10.0)
    assert list(g[1:, k]) == expect


@pytest.mark.parametrize("nhat,bin_expect", [
    ((0.0, 1.0), 50), ((1.0, 0.0), 75), ((-1.0, 0.0), 25), ((0.0, -1.0), 99),
    ((math.sqrt(0.5), math.sqrt(0.5)), 62),
])
def test_indenter_projection_and_bin(nhat, bin_expect):
    # node at distance 0.1 from the indenter centre along nhat, moving into it
    p = make_params(Gravity=0.0, IndVelocity=0.3)
    ix, iy = 8, 8
    cx = ix * p.cellsize - 0.1 * nhat[0]
    cy = iy * p.cellsize - 0.1 * nhat[1]
    k = ix + iy * p.GridX
    m = 1.5
    vin = np.array([0.3, 0.0]) - 2.0 * np.array(nhat) + 0.7 * np.array([-nhat[1], nhat[0]])
    g = node_grid(p, k, m, m * vin)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    fs = oracle.update_nodes(p, g, bins, cx, cy)
    vout = g[1:, k]
    n = np.array(nhat)
    # no relative normal velocity left, tangential part kept
    assert abs((vout - [0.3, 0.0]) @ n) < 1e-14
    assert (vout - [0.3, 0.0]) @ [-n[1], n[0]] == pytest.approx(0.7, rel=1e-13)
    f = (vin - vout) * m / p.InitialTimeStep
    assert np.allclose(fs, f, rtol=1e-12)
    nz = np.flatnonzero(bins)
    assert set(nz // 2) == {bin_expect}
    assert np.allclose(bins[2 * bin_expect:2 * bin_expect + 2], f, rtol=1e-12)


def test_indenter_separating_node_is_free():
    p = make_params(Gravity=0.0, IndVelocity=0.3)
    k = 8 + 8 * p.GridX
    cx, cy = 8 * p.cellsize - 0.1, 8 * p.cellsize
    g = node_grid(p, k, 1.0, [1.0, 0.2])        # moving away faster than the disc
    bins = np.zeros(200)
    fs = oracle.update_nodes(p, g, bins, cx, cy)
    assert list(g[1:, k]) == [1.0, 0.2] and not bins.any() and not fs.any()


# ---- G2P -------------------------------------------------------------------------------
def test_g2p_uniform_field():
    p = make_params()
    rng = np.random.default_rng(8)
    x = rng.uniform(0.4, 1.2, size=(30, 2))
    pts = make_points(p, x)
    grid = oracle.new_grid(p)
    grid[1, :p.GridTotal] = 0.37
    grid[2, :p.GridTotal] = -0.21
    assert oracle.g2p(p, pts, grid, True) == 0
    assert np.allclose(pts[6, :30], 0.37, rtol=1e-14)
    assert np.allclose(pts[7, :30], -0.21, rtol=1e-14)
    assert np.abs(pts[12:16, :30]).max() < 1e-12
    assert np.allclose(pts[4:6, :30], x.T + p.InitialTimeStep * np.array([[0.37], [-0.21]]), rtol=1e-15)


def test_g2p_linear_field_recovers_gradient():
    # v(x) = G x + c on the nodes -> v_p = G x_p + c, Bp = G (quadratic MLS/APIC)
    p = make_params(dt=1e-7)
    G = np.array([[0.3, -0.2], [0.5, 0.1]])
    c = np.array([0.05, -0.02])
    ii, jj = np.meshgrid(np.arange(p.GridX), np.arange(p.GridY), indexing="xy")
    X = np.stack([ii.ravel() * p.cellsize, jj.ravel() * p.cellsize])
    V = G @ X + c[:, None]
    grid = oracle.new_grid(p)
    grid[1:, :p.GridTotal] = V
    rng = np.random.default_rng(9)
    x = rng.uniform(0.4, 1.2, size=(30, 2))
    pts = make_points(p, x)
    oracle.g2p(p, pts, grid, False)
    vp = (G @ x.T) + c[:, None]
    assert np.allclose(pts[6:8, :30], vp, rtol=1e-12, atol=1e-14)
    Bp = pts[12:16, :30].T.reshape(30, 2, 2)
    assert np.allclose(Bp, G, rtol=1e-10, atol=1e-12)
    # Fe = (I + dt Bp) Fe for the intact elastic points
    F = pts[8:12, :30].T.reshape(30, 2, 2)
    assert np.allclose(F, np.eye(2) + p.InitialTimeStep * Bp, rtol=0, atol=1e-15)


def test_g2p_record_pq_only_when_asked():
    p = make_params()
    pts = make_points(p, [[0.6, 0.6]], Fe=[[1.0001, 0.0, 0.0, 0.9999]])
    grid = oracle.new_grid(p)
    oracle.g2p(p, pts, grid, False)
    assert pts[1, 0] == 0 and pts[2, 0] == 0
    oracle.g2p(p, pts, grid, True)
    assert pts[2, 0] > 0
