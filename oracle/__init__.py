"""CPU oracle for the MLS-MPM substep -- TEST INFRASTRUCTURE ONLY.

Python bindings of oracle/liboracle.so (a C++ restatement of the reference
substep, see icempm_oracle.cpp) plus a numpy restatement of the raw-point
ingest (SnapshotManager::LoadRawPoints, snapshotmanager.cpp:186-286).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package, and only to check the HIP path or to time the CPU baseline.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_long, c_size_t, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
FAST_LIB_PATH = os.path.join(HERE, "liboracle_fast.so")   # the CPU-baseline timing build

_lib = None
_lib_fast = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    """The parity build (-O2, no contraction)."""
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH)
    return _lib


def lib_fast() -> ctypes.CDLL:
    """The timing build of the same source (-O3, AVX2), for the CPU baseline."""
    global _lib_fast
    if _lib_fast is None:
        _lib_fast = _load(FAST_LIB_PATH)
    return _lib_fast


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"oracle/{os.path.basename(path)} is not built (make -C oracle)")
    L = ctypes.CDLL(path)
    dp = POINTER(c_double)
    ip = POINTER(c_int)
    L.oracle_simparams_size.restype = c_int
    L.oracle_simparams_offsets.argtypes = [POINTER(c_long), c_int]
    L.oracle_simparams_offsets.restype = c_int
    L.oracle_svd2x2.argtypes = [dp, dp, dp, dp]
    L.oracle_kirchhoff.argtypes = [dp, c_double, c_double, dp]
    L.oracle_compute_pq.argtypes = [dp, c_double, c_double, dp]
    L.oracle_grain_params.argtypes = [c_void_p, c_int, dp]
    L.oracle_failure_check.argtypes = [c_void_p, c_int, c_double, c_double]
    L.oracle_failure_check.restype = c_int
    L.oracle_dp_return.argtypes = [c_void_p, dp, dp]
    L.oracle_reset_grid.argtypes = [c_void_p, dp]
    L.oracle_p2g.argtypes = [c_void_p, dp, dp]
    L.oracle_p2g.restype = c_int
    L.oracle_update_nodes.argtypes = [c_void_p, dp, dp, c_double, c_double, dp]
    L.oracle_g2p.argtypes = [c_void_p, dp, dp, c_int]
    L.oracle_g2p.restype = c_int
    L.oracle_run.argtypes = [c_void_p, dp, dp, dp, c_int, dp, c_int]
    L.oracle_run.restype = c_int
    L.oracle_cell_keys_f32.argtypes = [POINTER(c_float), c_int, c_double, c_int, ip]
    L.oracle_cell_keys_soa.argtypes = [dp, c_size_t, c_int, c_double, c_int, ip]
    L.oracle_stable_order.argtypes = [ip, c_int, ip]
    L.oracle_mlp_loss_grad.argtypes = [dp, dp, dp, c_int, dp, dp]
    L.oracle_mlp_loss_grad.restype = c_double
    return L


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(POINTER(c_double))


def _ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(POINTER(c_int))


# -- 2x2 algebra and constitutive model ---------------------------------------
def svd2x2(F) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(4))
    U, S, V = np.empty(4), np.empty(2), np.empty(4)
    lib().oracle_svd2x2(_dp(F), _dp(U), _dp(S), _dp(V))
    return U.reshape(2, 2), S, V.reshape(2, 2)


def kirchhoff(F, mu: float, kappa: float) -> np.ndarray:
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(4))
    tau = np.empty(4)
    lib().oracle_kirchhoff(_dp(F), mu, kappa, _dp(tau))
    return tau.reshape(2, 2)


def compute_pq(F, mu: float, kappa: float) -> dict:
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(4))
    out = np.empty(16)
    lib().oracle_compute_pq(_dp(F), mu, kappa, _dp(out))
    return {"U": out[0:4].reshape(2, 2), "S": out[4:6].copy(), "V": out[6:10].reshape(2, 2),
            "Je": out[10], "p": out[11], "q": out[12], "shat": out[13:15].copy()}


def grain_params(prms, grain: int) -> dict:
    out = np.empty(5)
    lib().oracle_grain_params(ctypes.byref(prms), grain, _dp(out))
    return dict(zip(("pmin", "pmax", "qmax", "beta", "Msq"), out.tolist()))


def failure_check(prms, grain: int, p: float, q: float) -> int:
    return lib().oracle_failure_check(ctypes.byref(prms), grain, p, q)


def dp_return(prms, F, Jp_inv: float) -> tuple[np.ndarray, float]:
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(4)).copy()
    J = np.array([Jp_inv], dtype=np.float64)
    lib().oracle_dp_return(ctypes.byref(prms), _dp(F), _dp(J))
    return F.reshape(2, 2), float(J[0])


# -- substep pieces on SoA buffers ---------------------------------------------
def _check_buffers(prms, pts=None, grid=None, bins=None) -> None:
    """The C routines index by prms' sizes; refuse buffers smaller than that."""
    if grid is not None and (grid.shape != (3, prms.nGridPitch) or prms.nGridPitch < prms.GridTotal):
        raise ValueError("grid must be (3, nGridPitch) with nGridPitch >= GridTotal")
    if pts is not None and (pts.shape != (16, prms.nPtsPitch) or prms.nPtsPitch < prms.nPts):
        raise ValueError("points must be (16, nPtsPitch) with nPtsPitch >= nPts")
    if bins is not None and bins.size < 2 * prms.n_indenter_subdivisions:
        raise ValueError("bins must hold 2 * n_indenter_subdivisions doubles")


def new_grid(prms) -> np.ndarray:
    if prms.nGridPitch < prms.GridTotal:
        raise ValueError("set prms.nGridPitch (>= GridTotal) first")
    return np.zeros((3, prms.nGridPitch), dtype=np.float64)


def p2g(prms, pts: np.ndarray, grid: np.ndarray) -> int:
    _check_buffers(prms, pts, grid)
    return lib().oracle_p2g(ctypes.byref(prms), _dp(pts), _dp(grid))


def update_nodes(prms, grid: np.ndarray, bins: np.ndarray, ix: float, iy: float) -> np.ndarray:
    _check_buffers(prms, None, grid, bins)
    fs = np.zeros(2)
    lib().oracle_update_nodes(ctypes.byref(prms), _dp(grid), _dp(bins), ix, iy, _dp(fs))
    return fs


def g2p(prms, pts: np.ndarray, grid: np.ndarray, record: bool) -> int:
    _check_buffers(prms, pts, grid)
    return lib().oracle_g2p(ctypes.byref(prms), _dp(pts), _dp(grid), 1 if record else 0)


def run(prms, pts: np.ndarray, grid: np.ndarray, bins: np.ndarray, nsub: int,
        nthreads: int = 1, fast: bool = False) -> tuple[np.ndarray, int]:
    """nsub substeps with Model::Step semantics (prms is advanced); fast
    selects the timing build."""
    _check_buffers(prms, pts, grid, bins)
    hist = np.zeros((max(nsub, 1), 2), dtype=np.float64)
    L = lib_fast() if fast else lib()
    err = L.oracle_run(ctypes.byref(prms), _dp(pts), _dp(grid), _dp(bins), nsub, _dp(hist), nthreads)
    return hist[:nsub], err


def cell_keys_f32(xy: np.ndarray, hinv: float, gridx: int) -> np.ndarray:
    xy = np.ascontiguousarray(xy, dtype=np.float32)
    keys = np.empty(xy.shape[0], dtype=np.int32)
    lib().oracle_cell_keys_f32(xy.ctypes.data_as(POINTER(c_float)), xy.shape[0], hinv, gridx,
                               _ip(keys))
    return keys


def cell_keys_soa(pts: np.ndarray, n: int, hinv: float, gridx: int) -> np.ndarray:
    keys = np.empty(n, dtype=np.int32)
    lib().oracle_cell_keys_soa(_dp(pts), pts.shape[1], n, hinv, gridx, _ip(keys))
    return keys


def stable_order(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    perm = np.empty(keys.shape[0], dtype=np.int32)
    lib().oracle_stable_order(_ip(keys), keys.shape[0], _ip(perm))
    return perm


# -- ingest (SnapshotManager::LoadRawPoints, snapshotmanager.cpp:186-286) -------
def ingest(prms, xy: np.ndarray, grains: np.ndarray, volume: float, pitch: int) -> np.ndarray:
    """Restatement of LoadRawPoints on arrays: sort by the raw cell key
    (stable), bounding box in float, block offsets, indenter placement,
    particle volume and mass, initial state.  Updates prms (nPts, Volume,
    bounds, indenter position, ParticleVolume/Mass, nPtsPitch) and returns
    the (16, pitch) point SoA."""
    xy = np.ascontiguousarray(xy, dtype=np.float32).copy()
    grains = np.ascontiguousarray(grains, dtype=np.int16).copy()
    n = xy.shape[0]
    prms.nPts = n
    if prms.SortPointsByGridCell:
        order = stable_order(cell_keys_f32(xy, prms.cellsize_inv, prms.GridX))
        xy = xy[order]
        grains = grains[order]
    prms.Volume = float(np.float32(volume))
    prms.xmin = float(xy[:, 0].min())
    prms.xmax = float(xy[:, 0].max())
    length = float(np.float32(prms.xmax - prms.xmin))
    prms.ymin = float(xy[:, 1].min())
    prms.ymax = float(xy[:, 1].max())
    h = prms.cellsize
    prms.nxmin = int(math.floor(prms.xmin / h))
    prms.nxmax = int(math.ceil(prms.xmax / h))
    prms.nymin = int(math.floor(prms.ymin / h))
    prms.nymax = int(math.ceil(prms.ymax / h))
    box_x = prms.GridX * h
    x_offset = (box_x - length) / 2
    y_offset = 2 * h
    block_left = x_offset
    block_top = prms.ymax + y_offset
    r = prms.IndDiameter / 2
    ht = r - prms.IndDepth
    x_ind_offset = math.sqrt(r * r - ht * ht)
    prms.indenter_x = math.floor((block_left - x_ind_offset) / h) * h
    if prms.SetupType == 0:
        prms.indenter_y = block_top + ht
    elif prms.SetupType == 1:
        prms.indenter_y = math.ceil(block_top / h) * h
    prms.indenter_x_initial = prms.indenter_x
    prms.indenter_y_initial = prms.indenter_y
    prms.ParticleVolume = prms.Volume / n
    prms.ParticleMass = prms.ParticleVolume * prms.Density
    prms.nPtsPitch = pitch

    pts = np.zeros((16, pitch), dtype=np.float64)
    u8 = pts[0].view(np.uint8)
    u8[:n] = 0
    u8[pitch:pitch + 2 * n].view(np.int16)[:] = grains
    pts[3, :n] = 1.0
    # float32 += float64 rounds the sum back to float32
    pts[4, :n] = (xy[:, 0].astype(np.float64) + x_offset).astype(np.float32).astype(np.float64)
    pts[5, :n] = (xy[:, 1].astype(np.float64) + y_offset).astype(np.float32).astype(np.float64)
    pts[8, :n] = 1.0
    pts[11, :n] = 1.0
    return pts


# -- learned return map --------------------------------------------------------
MLP_PARAMS = 6288


def mlp_loss_grad(params, X, T, want_grad=True):
    """FP64 loss, gradient and outputs of the MLP (layout as in csrc/mlp.hip)."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    B = X.shape[0]
    g = np.zeros(MLP_PARAMS) if want_grad else None
    Y = np.zeros((B, 3))
    L = lib().oracle_mlp_loss_grad(_dp(P), _dp(X), _dp(T), B, _dp(g) if want_grad else None, _dp(Y))
    return L, g, Y
