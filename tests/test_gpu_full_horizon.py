"""The plain configurations over their stated horizons, against golden
vectors the CPU oracle wrote (tests/golden/make_full_horizon.py).

  indent-60k, 20,000 substeps = 8 Model::Step cycles (model.cpp:11-47)
    ORDERED   bit-exact with the sequential oracle: every substep's indenter
              force, every cycle's Indenter_2D bins, the final point SoA.
    FAST      indenter force relative L2 per cycle, and the crushed flags
              against the oracle's every 500 substeps; the substep of the
              oracle's first crush and the first checkpoint whose flags
              differ are printed (the pre-fracture horizon).
  indent-1.6M, 500 substeps (FAST): force history and bins within 1e-6 of
              the oracle, crushed flags identical every 100 substeps.

Both start by showing that the device ingest produced the oracle's input
bits (SHA-256 of the point SoA).
"""
import hashlib
import os

import numpy as np
import pytest

from icempm_amd import INDENT_1M6, INDENT_60K, MODE_FAST, MODE_ORDERED, Model

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / n if n > 0 else np.linalg.norm(a - b)


def gold(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated (python tests/golden/make_full_horizon.py)")
    return np.load(path)


def crushed_at(fx, k):
    off = fx["crushed_off"]
    return fx["crushed_idx"][off[k]:off[k + 1]]


def test_indent60k_20000_substeps_ordered_bit_exact(gpu_available):
    fx = gold("indent60k_20000.npz")
    case = INDENT_60K
    m = Model(MODE_ORDERED)
    try:
        m.load_case(case)
        assert sha(m.points()) == str(fx["init_sha256"])
        N = m.prms.UpdateEveryNthStep
        assert N == 2500 and fx["hist"].shape[0] == 8 * N
        for cyc in range(8):
            say("indent-60k ORDERED cycle", cyc + 1, "of 8")
            assert m.step()
            m.wait_cycle()
            h = m.last_force_history(N)
            ref = fx["hist"][cyc * N:(cyc + 1) * N]
            diff = np.flatnonzero((h != ref).any(axis=1))
            assert diff.size == 0, f"cycle {cyc}: first differing substep {cyc * N + diff[0]}"
            assert np.array_equal(m.host_force(), fx["bins"][cyc]), cyc
        assert m.prms.SimulationStep == 20000
        assert m.prms.SimulationTime == float(fx["simulation_time"])
        assert sha(m.host_points()) == str(fx["final_sha256"])
    finally:
        m.close()


def test_indent60k_20000_substeps_fast(gpu_available):
    fx = gold("indent60k_20000.npz")
    case = INDENT_60K
    n = case.n_points
    ck = int(fx["checkpoint"])
    m = Model(MODE_FAST)
    try:
        m.load_case(case)
        assert sha(m.points()) == str(fx["init_sha256"])
        hist = np.zeros_like(fx["hist"])
        first_mismatch = None
        for k in range(fx["hist"].shape[0] // ck):
            hist[k * ck:(k + 1) * ck] = m.run(ck)
            got = np.flatnonzero(m.points()[0].view(np.uint8)[:n])
            if first_mismatch is None and not np.array_equal(got, crushed_at(fx, k)):
                first_mismatch = (k + 1) * ck
        assert m.error_flag() == 0
        pos = m.points()[4:6, :n:int(fx["stride"])]
    finally:
        m.close()
    counts = fx["crushed_count"]
    first_crush = int(np.argmax(counts > 0)) + 1 if counts.any() else None
    N = 2500
    per_cycle = [rel_l2(hist[c * N:(c + 1) * N], fx["hist"][c * N:(c + 1) * N]) for c in range(8)]
    say("indent-60k FAST: force rel L2 per cycle", ["%.2e" % e for e in per_cycle])
    say("indent-60k FAST: oracle's first crush at substep", first_crush,
        "; crushed flags first differ at checkpoint", first_mismatch,
        "; crushed at the end: oracle", int(counts[-1]))
    say("indent-60k FAST: final positions max difference %.2e m" % np.abs(pos - fx["pos_final"]).max())
    # before the flags part, the run is held to the FP64 bar
    horizon = first_mismatch if first_mismatch is not None else hist.shape[0]
    pre = (horizon // N) * N if horizon >= N else horizon - horizon % ck
    if pre > 0:
        assert rel_l2(hist[:pre], fx["hist"][:pre]) <= 1e-6
    assert np.isfinite(hist).all()


def test_indent1m6_500_substeps_fast(gpu_available):
    fx = gold("indent1.6M_500.npz")
    case = INDENT_1M6
    n = case.n_points
    ck = int(fx["checkpoint"])
    m = Model(MODE_FAST)
    try:
        m.load_case(case)
        assert sha(m.points()) == str(fx["init_sha256"])
        hist = np.zeros_like(fx["hist"])
        for k in range(fx["hist"].shape[0] // ck):
            say("indent-1.6M FAST substeps", k * ck, "to", (k + 1) * ck)
            hist[k * ck:(k + 1) * ck] = m.run(ck, reset_bins_at_cycle=True)
            got = np.flatnonzero(m.points()[0].view(np.uint8)[:n])
            assert np.array_equal(got, crushed_at(fx, k)), (k + 1) * ck
        assert m.error_flag() == 0
        bins = m.bins()
        pos = m.points()[4:6, :n:int(fx["stride"])]
    finally:
        m.close()
    e = rel_l2(hist, fx["hist"])
    eb = rel_l2(bins, fx["bins"][-1])
    dx = np.abs(pos - fx["pos_final"]).max()
    say(f"indent-1.6M FAST 500 substeps: force rel L2 {e:.2e}, bins {eb:.2e}, positions {dx:.2e} m, "
        f"crushed {int(fx['crushed_count'][-1])}")
    assert e <= 1e-6 and eb <= 1e-6
    assert dx <= 1e-12
