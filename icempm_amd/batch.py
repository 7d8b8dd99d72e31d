"""Batches of independent indentation cases on one GPU (C ABI
``icempm_batch_*``): many cases of uneven size advanced together, one grid
update and one particle-pass launch per substep for all of them, the substep
loop captured in a HIP graph.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import POINTER, byref, c_double, c_float, c_int, c_uint64, c_void_p

import numpy as np

from . import _lib
from ._lib import (BATCH_F32, BATCH_F64, BATCH_TUNE_CHUNK, BATCH_TUNE_MAP_MIN_ACCEPT, BATCH_TUNE_RESORT,
                   MAP_EXACT, MAP_HARVEST, MAP_LEARNED, MAP_SHEAR, MAP_TENSION, IcempmError, SimParams)
from .model import MODE_FAST, IndentationCase, Model, nacc_strengths


def batch_cases(n_cases: int, n_min: int, n_max: int, seed: int, h: float = 0.012,
                dt: float = 2e-6, e_range=(5e9, 10e9), tensile_range=(0.5e6, 2e6),
                speed_range=(0.1, 1.0)) -> list[IndentationCase]:
    """Seeded cases: point counts log-uniform in [n_min, n_max] (4 points per
    cell, block aspect 8:3 like the 2.4 m x 0.9 m block), Young's modulus,
    tensile strength (NACC M = 1.2, beta = 0.8 kept) and disc speed uniform in
    their ranges; grid margins as in the 60k case (56 columns, 53 rows)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_cases):
        n = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        cells = n / 4
        ncy = max(8, int(round(math.sqrt(cells * 3 / 8))))
        ncx = max(16, int(round(cells / ncy)))
        E = rng.uniform(*e_range)
        T = rng.uniform(*tensile_range)
        V = rng.uniform(*speed_range)
        C, T2, S = nacc_strengths(1.2, 0.8, T)
        out.append(IndentationCase(
            name=f"batch{seed}-{i}", ncx=ncx, ncy=ncy, per_side=2, h=h, grid_x=ncx + 56,
            grid_y=ncy + 53, dt=dt, steps=2000, youngs=E, compressive=C, tensile=T2, shear=S,
            ind_diameter=0.3, ind_velocity=V, ind_depth=0.1, seed=0x1CE5EED + i))
    return out


class Batch:
    def __init__(self, precision: int = BATCH_F64):
        self._H = _lib.hip()
        b = c_void_p()
        rc = self._H.icempm_batch_create(byref(b), precision)
        if rc != 0:
            raise IcempmError(rc, "icempm_batch_create failed")
        self._b = b
        self.precision = precision
        self.cases: list[IndentationCase] = []
        self.initial: list[tuple[SimParams, np.ndarray]] = []

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self._H.icempm_batch_last_error(self._b) or b""
            raise IcempmError(rc, msg.decode(errors="replace"))

    def close(self) -> None:
        if getattr(self, "_b", None):
            self._H.icempm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, model: Model, case: IndentationCase) -> int:
        """Ingest the case through the model (LoadRawPoints) and add its state."""
        model.load_case(case)
        soa = np.ascontiguousarray(model.points())
        return self.add_state(case, model.prms.copy(), soa)

    def add_state(self, case: IndentationCase, prms: SimParams, soa: np.ndarray) -> int:
        """Add a case from its parameters and point SoA (16 x pitch doubles)."""
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        p = prms.copy()
        idx = self._H.icempm_batch_add_case(self._b, byref(p), ctypes.sizeof(SimParams),
                                            soa.ctypes.data_as(POINTER(c_double)), soa.shape[1])
        if idx < 0:
            self._check(-idx)
        self.cases.append(case)
        self.initial.append((p, soa))
        return idx

    def finalize(self) -> None:
        self._check(self._H.icempm_batch_finalize(self._b))

    def prepare(self, n: int) -> None:
        """Capture, instantiate and upload the graphs of an n-substep run ahead
        of time (icempm_batch_prepare)."""
        self._check(self._H.icempm_batch_prepare(self._b, n))

    def run(self, n: int, use_graph: bool = True) -> float:
        """n substeps of every case; returns the device time in ms."""
        self._check(self._H.icempm_batch_run(self._b, n, 1 if use_graph else 0))
        ms = c_float(0)
        self._check(self._H.icempm_batch_last_ms(self._b, byref(ms)))
        return ms.value

    @property
    def num_points(self) -> int:
        return int(self._H.icempm_batch_num_points(self._b))

    @property
    def num_nodes(self) -> int:
        return int(self._H.icempm_batch_num_nodes(self._b))

    def kernel_timing(self, stride: int) -> None:
        """Time every stride-th fused particle pass of stream runs (0: off)."""
        self._check(self._H.icempm_batch_kernel_timing(self._b, int(stride)))

    def kernel_timing_read(self) -> tuple[float, int]:
        ms, n = c_double(0), c_int(0)
        self._check(self._H.icempm_batch_kernel_timing_read(self._b, byref(ms), byref(n)))
        return ms.value, n.value

    def force_history(self, c: int, n: int) -> np.ndarray:
        out = np.zeros((n, 2))
        self._check(self._H.icempm_batch_force_history(self._b, c, out.ctypes.data_as(POINTER(c_double)), n))
        return out

    def force_bins(self, c: int) -> np.ndarray:
        """Case c's indenter force bins (Indenter_2D: fx, fy per angular bin)
        summed over the last run."""
        nb = int(self.initial[c][0].n_indenter_subdivisions)
        out = np.zeros(2 * nb)
        self._check(self._H.icempm_batch_force_bins(self._b, c, out.ctypes.data_as(POINTER(c_double)), nb))
        return out

    def points(self, c: int) -> np.ndarray:
        pitch = self.initial[c][1].shape[1]
        out = np.zeros((16, pitch))
        self._check(self._H.icempm_batch_read_points(self._b, c, out.ctypes.data_as(POINTER(c_double)), pitch))
        return out

    def error_flag(self) -> int:
        f = c_int(0)
        self._check(self._H.icempm_batch_error_flag(self._b, byref(f)))
        return f.value

    # -- tuning, re-sort, checkpoint -------------------------------------------
    def set_tuning(self, chunk: int | None = None, resort_every: int | None = None) -> None:
        """Points per chunk (64, 128, 256) and the re-sort cadence in substeps (0: never)."""
        if chunk is not None:
            self._check(self._H.icempm_batch_set_tuning(self._b, BATCH_TUNE_CHUNK, int(chunk)))
        if resort_every is not None:
            self._check(self._H.icempm_batch_set_tuning(self._b, BATCH_TUNE_RESORT, int(resort_every)))

    def tuning(self) -> tuple[int, int]:
        return (int(self._H.icempm_batch_get_tuning(self._b, BATCH_TUNE_CHUNK)),
                int(self._H.icempm_batch_get_tuning(self._b, BATCH_TUNE_RESORT)))

    def resort(self) -> None:
        self._check(self._H.icempm_batch_resort(self._b))

    def checkpoint(self) -> None:
        self._check(self._H.icempm_batch_checkpoint(self._b))

    def restore(self) -> None:
        self._check(self._H.icempm_batch_restore(self._b))

    def autotune(self, probe_steps: int = 200) -> list[tuple[int, int, float]]:
        """Times every (points per chunk, re-sort cadence) candidate over
        probe_steps substeps from the current state, keeps the fastest and
        returns the state to where it was.  Rows: (chunk, re-sort, ms/substep)."""
        rep = (c_double * 64)()
        rows = c_int(0)
        self._check(self._H.icempm_batch_autotune(self._b, int(probe_steps), rep, 21, byref(rows)))
        return [(int(rep[3 * i]), int(rep[3 * i + 1]), float(rep[3 * i + 2])) for i in range(rows.value)]

    # -- return map of crushed points ------------------------------------------
    def set_exact(self) -> None:
        self._check(self._H.icempm_batch_set_return_map(self._b, MAP_EXACT, None, None, 0.0, 0))

    def set_harvest(self, capacity: int) -> None:
        """Exact map, recording up to ``capacity`` (features, targets) pairs."""
        self._check(self._H.icempm_batch_set_return_map(self._b, MAP_HARVEST, None, None, 0.0,
                                                        int(capacity)))
        self._hcap = int(capacity)

    def set_learned(self, params: np.ndarray, norm: np.ndarray, tol: float = 1e-4) -> None:
        """MLP proposal accepted where the yield residual is <= tol, else the
        exact map.  norm: 16 floats (input mean 5, input scale 5, output mean
        3, output scale 3)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        if p.shape != (6288,) or nm.shape != (16,):
            raise ValueError("params must have 6288 and norm 16 entries")
        fp = POINTER(c_float)
        self._check(self._H.icempm_batch_set_return_map(self._b, MAP_LEARNED, p.ctypes.data_as(fp),
                                                        nm.ctypes.data_as(fp), float(tol), 0))

    def read_harvest(self, cap: int) -> tuple[np.ndarray, np.ndarray, int]:
        """(X (n x 5), T (n x 3), pairs offered); n <= cap and <= the capacity."""
        X = np.zeros((max(cap, 0), 5), dtype=np.float32)
        T = np.zeros((max(cap, 0), 3), dtype=np.float32)
        tot = c_uint64(0)
        fp = POINTER(c_float)
        self._check(self._H.icempm_batch_read_harvest(self._b, X.ctypes.data_as(fp), T.ctypes.data_as(fp),
                                                      int(cap), byref(tot)))
        n = min(int(tot.value), cap, getattr(self, "_hcap", 0))
        return X[:n], T[:n], int(tot.value)

    def set_map_branches(self, mask: int) -> None:
        """Branches the network proposes for (MAP_TENSION | MAP_SHEAR); the
        others take the exact map in the particle pass."""
        self._check(self._H.icempm_batch_set_map_branches(self._b, int(mask)))

    def map_branches(self) -> int:
        return int(self._H.icempm_batch_get_map_branches(self._b))

    def set_map_min_accept(self, percent: int) -> None:
        """Acceptance threshold below which autotune switches a branch off."""
        self._check(self._H.icempm_batch_set_tuning(self._b, BATCH_TUNE_MAP_MIN_ACCEPT, int(percent)))

    def map_branch_stats(self, reset: bool = False) -> tuple[int, int, int, int]:
        """(tension accepted, tension fell back, shear accepted, shear fell back)"""
        out = (c_uint64 * 4)()
        self._check(self._H.icempm_batch_map_branch_stats(self._b, out, 1 if reset else 0))
        return tuple(int(x) for x in out)

    def autotune_branch_stats(self) -> tuple[int, int, int, int]:
        """The last autotune's per-branch counts over its probes (as map_branch_stats)."""
        out = (c_uint64 * 4)()
        self._check(self._H.icempm_batch_autotune_branch_stats(self._b, out))
        return tuple(int(x) for x in out)

    def map_stats(self, reset: bool = False) -> tuple[int, int, int]:
        """(points accepted from the MLP, points that fell back to the exact
        map, crushed points inside the surface), summed over substeps."""
        out = (c_uint64 * 3)()
        self._check(self._H.icempm_batch_map_stats(self._b, out, 1 if reset else 0))
        return int(out[0]), int(out[1]), int(out[2])


def ingest_cases(cases: list[IndentationCase]) -> list[tuple[SimParams, np.ndarray]]:
    """Each case's parameters and point SoA after LoadRawPoints on the GPU."""
    out = []
    m = Model(MODE_FAST)
    try:
        for c in cases:
            m.load_case(c)
            out.append((m.prms.copy(), np.ascontiguousarray(m.points())))
    finally:
        m.close()
    return out


def build_batch(cases: list[IndentationCase], precision: int = BATCH_F64,
                states: list[tuple[SimParams, np.ndarray]] | None = None) -> Batch:
    """A finalized batch of the cases (ingested here unless states are given)."""
    if states is None:
        states = ingest_cases(cases)
    b = Batch(precision)
    for c, (p, soa) in zip(cases, states):
        b.add_state(c, p, soa)
    b.finalize()
    return b
