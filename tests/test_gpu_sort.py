"""Cell-key sort on the device (replaces SnapshotManager::SortPoints,
snapshotmanager.cpp:288-304): keys and order must equal the oracle's stable
order bit for bit."""
import numpy as np
import pytest

import oracle
from icempm_amd import INDENT_60K, MODE_FAST, Model, configure_params, lattice_points

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def model(gpu_available):
    m = Model(MODE_FAST)
    yield m
    m.close()


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (1000, 2), (2048, 3), (2049, 4), (100003, 5)])
def test_sort_raw_points_matches_stable_order(model, n, seed):
    say("sort raw", n)
    rng = np.random.default_rng(seed)
    hinv, gx = 1 / 0.012, 256
    # a coarse lattice gives many equal keys; a few negative coordinates give
    # negative keys
    xy = (rng.integers(-20, 3000, size=(n, 2)) * 0.0011).astype(np.float32)
    g = rng.integers(0, 1000, size=n).astype(np.int16)
    sx, sg = model.sort_raw_points(xy, g, hinv, gx)
    order = oracle.stable_order(oracle.cell_keys_f32(xy, hinv, gx))
    assert np.array_equal(sx, xy[order])
    assert np.array_equal(sg, g[order])


def permute_soa(buf, n, order):
    out = buf.copy()
    pitch = buf.shape[1]
    out[1:, :n] = buf[1:, order]
    u8o, u8 = out[0].view(np.uint8), buf[0].view(np.uint8)
    u8o[:n] = u8[:n][order]
    u8o[pitch:pitch + 2 * n].view(np.int16)[:] = u8[pitch:pitch + 2 * n].view(np.int16)[order]
    u8o[3 * pitch:3 * pitch + n] = u8[3 * pitch:3 * pitch + n][order]
    return out


def test_sort_points_soa(model):
    say("sort soa")
    case = INDENT_60K
    configure_params(model.prms, case)
    xy, g = lattice_points(case)
    model.load_raw_points(xy, g, case.volume)
    model.run(20)                                   # non-trivial state in every row
    n = case.n_points
    buf = model.points()
    rng = np.random.default_rng(9)
    shuffle = rng.permutation(n)
    shuffled = permute_soa(buf, n, shuffle)
    shuffled[0].view(np.uint8)[3 * buf.shape[1]:3 * buf.shape[1] + n] = rng.integers(0, 255, n)
    model.host_points()[:] = shuffled
    from icempm_amd import _lib
    _lib.check_ctx(_lib.hip().icempm_transfer_points_to_device(model.ctx), model.ctx)
    perm, keys = model.sort_points()
    p = model.prms
    ref_keys = oracle.cell_keys_soa(shuffled, n, p.cellsize_inv, p.GridX)
    order = oracle.stable_order(ref_keys)
    assert np.array_equal(perm, order)
    assert np.array_equal(keys, ref_keys[order])
    assert np.array_equal(model.points().view(np.uint64), permute_soa(shuffled, n, order).view(np.uint64))
    # the sorted state keeps stepping
    model.run(5)
    assert model.error_flag() == 0
