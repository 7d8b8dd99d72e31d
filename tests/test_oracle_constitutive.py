"""Pins the oracle's constitutive routines (gpu_implementation5.cu:24-141)
with closed forms, invariants and cross-checks between reference formulas."""
import math

import numpy as np
import pytest

import oracle
from icempm_amd import SimParams, host, nacc_strengths

MU, KAPPA = 3.4615384615384617e9, 7.5e9


def rot(th):
    return np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])


def params(**kw):
    p = SimParams()
    host().icempm_simparams_reset(__import__("ctypes").byref(p))
    C, T, S = nacc_strengths(1.2, 0.8, 1e6)
    p.YoungsModulus = 9e9
    p.PoissonsRatio = 0.3
    p.IceCompressiveStrength, p.IceTensileStrength, p.IceShearStrength = C, T, S
    p.GrainVariability = 0.0
    p.DP_tan_phi = math.tan(70 * math.pi / 180)
    p.DP_threshold_p = 1e3
    for k, v in kw.items():
        setattr(p, k, v)
    host().icempm_simparams_derive(__import__("ctypes").byref(p))
    return p


# ---- Kirchhoff stress --------------------------------------------------------
def test_kirchhoff_identity_is_stress_free():
    assert np.array_equal(oracle.kirchhoff(np.eye(2), MU, KAPPA), np.zeros((2, 2)))


@pytest.mark.parametrize("s", [0.9, 0.999, 1.0001, 1.05])
def test_kirchhoff_isotropic(s):
    # F = sI: J = s^2, dev(FF^T) = 0 -> tau = kappa (s^4 - 1) I
    tau = oracle.kirchhoff(s * np.eye(2), MU, KAPPA)
    expect = KAPPA * (s ** 4 - 1)
    assert tau[0, 1] == 0 and tau[1, 0] == 0
    assert tau[0, 0] == pytest.approx(expect, rel=1e-12)
    assert tau[1, 1] == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("g", [1e-4, 0.01, 0.3])
def test_kirchhoff_simple_shear(g):
    # F = [[1, g], [0, 1]]: J = 1, b = [[1+g^2, g], [g, 1]],
    # dev(b) = [[g^2/2, g], [g, -g^2/2]] -> tau = mu dev(b)
    tau = oracle.kirchhoff(np.array([[1.0, g], [0.0, 1.0]]), MU, KAPPA)
    expect = MU * np.array([[g * g / 2, g], [g, -g * g / 2]])
    assert np.allclose(tau, expect, rtol=1e-12, atol=MU * 1e-15)


def test_kirchhoff_is_objective():
    # tau(R F) = R tau(F) R^T; fails if F^T F were used instead of F F^T
    rng = np.random.default_rng(3)
    for _ in range(200):
        F = np.eye(2) + 0.05 * rng.normal(size=(2, 2))
        R = rot(rng.uniform(-3, 3))
        lhs = oracle.kirchhoff(R @ F, MU, KAPPA)
        rhs = R @ oracle.kirchhoff(F, MU, KAPPA) @ R.T
        assert np.abs(lhs - rhs).max() <= 1e-6 * max(1.0, np.abs(rhs).max()) * 1e-3


# ---- trial state ---------------------------------------------------------------
def test_pq_identity():
    r = oracle.compute_pq(np.eye(2), MU, KAPPA)
    assert r["p"] == 0 and r["q"] == 0 and r["Je"] == 1


@pytest.mark.parametrize("a,b", [(1.01, 0.99), (0.98, 0.97), (1.0, 1.002), (1.1, -0.9)])
def test_pq_diagonal_closed_form(a, b):
    r = oracle.compute_pq(np.diag([a, b]), MU, KAPPA)
    J = a * b
    assert r["Je"] == pytest.approx(J, rel=1e-15)
    assert r["p"] == pytest.approx(-(KAPPA / 2) * (J * J - 1), rel=1e-9, abs=1e-3)
    assert r["q"] == pytest.approx(MU * abs(a * a - b * b) / abs(J), rel=1e-12)
    # q ignores rotations on either side
    R1, R2 = rot(0.3), rot(-1.1)
    r2 = oracle.compute_pq(R1 @ np.diag([a, b]) @ R2, MU, KAPPA)
    assert r2["q"] == pytest.approx(r["q"], rel=1e-9)
    assert r2["p"] == pytest.approx(r["p"], rel=1e-9, abs=1e-3)


def test_pressure_is_half_the_stress_trace_convention():
    # The stress uses kappa (J^2-1) while the trial pressure uses kappa/2
    # (gpu_implementation5.cu:32 vs :40); both are kept as written.
    F = 0.999 * np.eye(2)
    tau = oracle.kirchhoff(F, MU, KAPPA)
    p = oracle.compute_pq(F, MU, KAPPA)["p"]
    assert p == pytest.approx(-0.5 * tau[0, 0], rel=1e-12)


# ---- per-grain strengths and the NACC surface ----------------------------------
def test_grain_params_uniform_material_matches_nacc_constants():
    p = params()
    g = oracle.grain_params(p, 0)
    assert g["pmax"] == p.IceCompressiveStrength
    assert g["pmin"] == -p.IceTensileStrength
    assert g["qmax"] == p.IceShearStrength
    assert g["beta"] == pytest.approx(0.8, rel=1e-15)
    # GetParametersForGrain's M^2 equals ComputeCamClayParams2's NACC_Msq
    assert g["Msq"] == pytest.approx(p.NACC_Msq, rel=1e-14)
    assert math.sqrt(g["Msq"]) == pytest.approx(1.2, rel=1e-12)


def test_grain_msq_equals_camclay_msq_of_the_grain_strengths():
    """For every grain id 0..999 at GrainVariability 0.5, GetParametersForGrain's
    M^2 = 4 q^2 (1+2b) / (p (1+b))^2 (gpu_implementation5.cu:107) equals the
    square of ComputeCamClayParams2's M = 2 q sqrt(1+2b) / (p (1+b))
    (parameters_sim.cpp:141-146, run by the host library) evaluated on that
    grain's (pmax, -pmin, qmax); and beta = -pmin / pmax."""
    p = params(GrainVariability=0.5)
    for grain in range(1000):
        g = oracle.grain_params(p, grain)
        q = params(IceCompressiveStrength=g["pmax"], IceTensileStrength=-g["pmin"],
                   IceShearStrength=g["qmax"])
        assert g["beta"] == pytest.approx(q.NACC_beta, rel=1e-15), grain
        assert g["Msq"] == pytest.approx(q.NACC_Msq, rel=4e-15), grain


def test_grain_factors_are_periodic_with_21_30_11():
    """pmax repeats every 21 grain ids, pmin every 30, qmax every 11 (the
    moduli of GetParametersForGrain), and no shorter period fits."""
    p = params(GrainVariability=0.5)
    rows = [oracle.grain_params(p, g) for g in range(0, 700)]
    for key, period in (("pmax", 21), ("pmin", 30), ("qmax", 11)):
        v = np.array([r[key] for r in rows])
        assert np.array_equal(v[period:], v[:-period]), key
        for shorter in range(1, period):
            assert not np.array_equal(v[shorter:], v[:-shorter]), (key, shorter)


@pytest.mark.parametrize("gv", [0.5, 0.37])
def test_grain_factors_are_one_at_their_centres(gv):
    """Each factor is exactly 1 where its term -10 + g%21, -15 + (g+3)%30,
    -10 + (g+4)%11 vanishes: g = 10 (mod 21), 12 (mod 30), 6 (mod 11)."""
    p = params(GrainVariability=gv)
    for g in range(0, 1000):
        r = oracle.grain_params(p, g)
        if g % 21 == 10:
            assert r["pmax"] == p.IceCompressiveStrength, g
        else:
            assert r["pmax"] != p.IceCompressiveStrength, g
        if g % 30 == 12:
            assert r["pmin"] == -p.IceTensileStrength, g
        if g % 11 == 6:
            assert r["qmax"] == p.IceShearStrength, g


# Hand-computed factors at GrainVariability 0.5 (var = 1 + 0.5 * step * term):
#   g = 0    terms -10, -12, -6   -> 0.75,   0.802,  0.70
#   g = 37   37 % 21 = 16, 40 % 30 = 10, 41 % 11 = 8
#            terms   6,  -5, -2   -> 1.15,   0.9175, 0.90
#   g = -1   C++ % truncates: -1 % 21 = -1, 2 % 30 = 2, 3 % 11 = 3
#            terms -11, -13, -7   -> 0.725,  0.7855, 0.65
#            (a floored modulo would give 20, 2, 3 -> 1.25, 0.7855, 0.65)
#   g = -25  -25 % 21 = -4, -22 % 30 = -22, -21 % 11 = -10
#            terms -14, -37, -20  -> 0.65,   0.3895, 0.0
HAND_FACTORS = {0: (0.75, 0.802, 0.70), 37: (1.15, 0.9175, 0.90), -1: (0.725, 0.7855, 0.65),
                -25: (0.65, 0.3895, 0.0)}


@pytest.mark.parametrize("grain", sorted(HAND_FACTORS))
def test_grain_params_hand_values(grain):
    p = params(GrainVariability=0.5)
    r = oracle.grain_params(p, grain)
    f1, f2, f3 = HAND_FACTORS[grain]
    assert r["pmax"] == pytest.approx(p.IceCompressiveStrength * f1, rel=1e-14, abs=1e-6)
    assert r["pmin"] == pytest.approx(-p.IceTensileStrength * f2, rel=1e-14, abs=1e-6)
    assert r["qmax"] == pytest.approx(p.IceShearStrength * f3, rel=1e-14, abs=1e-6)


@pytest.mark.parametrize("grain", [0, 37, 101, 999, -1])
@pytest.mark.parametrize("frac", [0.1, 0.5, 0.9])
def test_failure_check_uses_the_grain_surface(grain, frac):
    """With GrainVariability 0.5 the NACC test follows the grain's own
    ellipse and tension line: points just inside and just outside the
    per-grain boundaries, computed from the grain's strengths by the
    closed forms q_b = sqrt(-M^2 (p + b pmax)(p - pmax) / (1 + 2b)) and
    q_b = q0 (p + 3e6) / 3e6 with q0 = 2 sqrt(-pmax pmin) qmax / (pmax - pmin)."""
    p = params(GrainVariability=0.5)
    r = oracle.grain_params(p, grain)
    pmax, pmin, qmax = r["pmax"], r["pmin"], r["qmax"]
    beta = -pmin / pmax
    M2 = (2 * qmax * math.sqrt(1 + 2 * beta) / (pmax * (1 + beta))) ** 2
    pp = frac * pmax
    qb = math.sqrt(-M2 * (pp + beta * pmax) * (pp - pmax) / (1 + 2 * beta))
    assert oracle.failure_check(p, grain, pp, qb * (1 - 1e-9)) == 0
    assert oracle.failure_check(p, grain, pp, qb * (1 + 1e-9)) == 1
    pt = -frac * 3e6
    q0 = 2 * math.sqrt(-pmax * pmin) * qmax / (pmax - pmin)
    qt = q0 * (pt + 3e6) / 3e6
    assert oracle.failure_check(p, grain, pt, qt * (1 - 1e-9)) == 0
    assert oracle.failure_check(p, grain, pt, qt * (1 + 1e-9) + 1e-9) == 1


def ellipse_q(p, g, pp):
    return math.sqrt(-g["Msq"] * (pp + g["beta"] * g["pmax"]) * (pp - g["pmax"]) / (1 + 2 * g["beta"]))


@pytest.mark.parametrize("frac", [0.05, 0.3, 0.6, 0.95])
def test_failure_ellipse_both_sides(frac):
    p = params()
    g = oracle.grain_params(p, 0)
    pp = frac * g["pmax"]
    qb = ellipse_q(p, g, pp)
    assert oracle.failure_check(p, 0, pp, qb * (1 - 1e-9)) == 0
    assert oracle.failure_check(p, 0, pp, qb * (1 + 1e-9)) == 1
    assert oracle.failure_check(p, 0, g["pmax"] * 1.001, 0.0) == 1


@pytest.mark.parametrize("pp", [-1e3, -5e5, -2.5e6])
def test_failure_tension_line_both_sides(pp):
    p = params()
    g = oracle.grain_params(p, 0)
    q0 = 2 * math.sqrt(-g["pmax"] * g["pmin"]) * g["qmax"] / (g["pmax"] - g["pmin"])
    ql = (q0 / 3e6) * (pp + 3e6)
    assert oracle.failure_check(p, 0, pp, ql * (1 - 1e-9)) == 0
    assert oracle.failure_check(p, 0, pp, ql * (1 + 1e-9)) == 1


def test_failure_surface_is_continuous_at_zero_pressure():
    # the tension line starts where the ellipse meets p = 0: q0 = M pmax sqrt(beta/(1+2beta))
    p = params()
    g = oracle.grain_params(p, 0)
    q0 = 2 * math.sqrt(-g["pmax"] * g["pmin"]) * g["qmax"] / (g["pmax"] - g["pmin"])
    assert ellipse_q(p, g, 0.0) == pytest.approx(q0, rel=1e-12)


def test_failure_hard_tension_cutoff():
    p = params()
    assert oracle.failure_check(p, 0, -3.0e6 * 1.0000001, 0.0) == 1
    assert oracle.failure_check(p, 0, -2.0e6, 0.0) == 0


# ---- Drucker-Prager return map ----------------------------------------------------
def _trial(F, p):
    return oracle.compute_pq(F, p.mu, p.kappa)


def test_dp_tension_branch_exact_volume():
    p = params()
    F = np.diag([1.0004, 1.0003])          # p_tr << -threshold (tension)
    r = _trial(F, p)
    assert r["p"] < -p.DP_threshold_p
    Fn, Jp = oracle.dp_return(p, F, 1.0)
    Jn = math.sqrt(1 + 2 * p.DP_threshold_p / p.kappa)
    assert np.linalg.det(Fn) == pytest.approx(Jn, rel=1e-14)
    assert np.allclose(Fn @ Fn.T, Jn * np.eye(2), atol=1e-15)
    assert Jp == pytest.approx(Jn / r["Je"], rel=1e-15)
    # the projected state sits at p = -threshold
    pn = -(p.kappa / 2) * (Jn * Jn - 1)
    assert pn == pytest.approx(-p.DP_threshold_p, rel=1e-6)


def test_dp_jp_below_one_takes_the_tension_branch():
    p = params()
    F = np.diag([0.9999, 0.9999])          # compression, but Jp_inv < 1
    Fn, Jp = oracle.dp_return(p, F, 0.99)
    Jn = math.sqrt(1 + 2 * p.DP_threshold_p / p.kappa)
    assert np.linalg.det(Fn) == pytest.approx(Jn, rel=1e-14)


def test_dp_cone_projection_invariants():
    p = params()
    rng = np.random.default_rng(7)
    n_proj = 0
    for _ in range(400):
        R1, R2 = rot(rng.uniform(-3, 3)), rot(rng.uniform(-3, 3))
        s0 = 1 - rng.uniform(0, 2e-4)
        s1 = s0 * (1 + rng.uniform(-3e-4, 3e-4))
        F = R1 @ np.diag([s0, s1]) @ R2
        r = _trial(F, p)
        if not (-p.DP_threshold_p <= r["p"] <= p.IceCompressiveStrength):
            continue
        qlim = min((r["p"] + p.DP_threshold_p) * p.DP_tan_phi, p.IceShearStrength)
        Fn, Jp = oracle.dp_return(p, F, 1.0)
        assert Jp == 1.0
        if r["q"] < qlim:
            assert np.array_equal(Fn, F)          # elastic: untouched
            continue
        n_proj += 1
        U, V = r["U"], r["V"]
        D = U.T @ Fn @ V                          # same rotations, new diagonal
        assert abs(D[0, 1]) < 1e-13 and abs(D[1, 0]) < 1e-13
        sn = np.array([D[0, 0], D[1, 1]])
        ss_tr = r["S"] ** 2
        # mean of sigma^2 is kept
        assert sn @ sn == pytest.approx(ss_tr.sum(), rel=1e-13)
        # deviatoric part measured with the trial J lands on q_lim
        dev = sn ** 2 - (sn ** 2).mean()
        assert math.sqrt(2) * p.mu / r["Je"] * np.linalg.norm(dev) == pytest.approx(qlim, rel=1e-6, abs=1e-3)
    assert n_proj > 20


def test_dp_zero_deviator_above_pmax():
    # p_tr > pmax -> q_lim = 0; an isotropic trial state (q_tr = 0) is projected
    # with the zero direction left at zero
    p = params(IceCompressiveStrength=1.0e3)
    F = 0.99999 * np.eye(2)
    r = _trial(F, p)
    assert r["p"] > p.IceCompressiveStrength and r["q"] == 0
    Fn, _ = oracle.dp_return(p, F, 1.0)
    s = math.sqrt(F[0, 0] ** 2)
    assert np.allclose(Fn, s * np.eye(2), rtol=0, atol=1e-15)
