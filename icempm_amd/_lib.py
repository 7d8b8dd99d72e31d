"""ctypes bindings of libicempm_hip.so (include/icempm.h) and
libicempm_host.so (include/icempm_host.h).

The libraries are loaded from this directory only; a missing library is an
error (there is no CPU fallback in the product).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_ulonglong, c_char_p, c_double, c_float, c_int, c_int16, c_long, c_size_t,
                    c_uint64, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(HERE, "libicempm_hip.so")
HOST_LIB_PATH = os.path.join(HERE, "libicempm_host.so")
H5_LIB_PATH = os.path.join(HERE, "libicempm_h5.so")

ICEMPM_OK = 0
ICEMPM_E_HIP = 1
ICEMPM_E_ARG = 2
ICEMPM_E_STATE = 3
ICEMPM_E_NODEV = 4
ICEMPM_E_IO = 5
ICEMPM_E_HOST = 6
MODE_FAST = 0
MODE_ORDERED = 1
TUNE_PASS_KIND = 0
BATCH_F64 = 0
BATCH_F32 = 1
MAP_EXACT = 0
MAP_HARVEST = 1
MAP_LEARNED = 2
BATCH_TUNE_CHUNK = 0
BATCH_TUNE_RESORT = 1
BATCH_TUNE_MAP_MIN_ACCEPT = 2
MAP_TENSION = 1
MAP_SHEAR = 2


class IcempmError(RuntimeError):
    """A nonzero status from the native library, with its message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


# ---------------------------------------------------------------------------
# icy::SimParams (parameters_sim.h:19-90 of the reference): same members in
# the same order, so the 432-byte layout matches the C++ struct.
_SIMPARAMS_FIELDS = [
    ("grid_array", c_void_p), ("indenter_force_accumulator", c_void_p), ("pts_array", c_void_p),
    ("nPtsPitch", c_size_t), ("nGridPitch", c_size_t),
    ("n_indenter_subdivisions", c_int),
    ("tpb_P2G", c_int), ("tpb_Upd", c_int), ("tpb_G2P", c_int),
    ("nPts", c_int),
    ("GridX", c_int), ("GridY", c_int), ("GridTotal", c_int),
    ("GridXDimension", c_double),
    ("InitialTimeStep", c_double), ("SimulationEndTime", c_double),
    ("Gravity", c_double), ("Density", c_double), ("PoissonsRatio", c_double),
    ("YoungsModulus", c_double),
    ("lambda_", c_double), ("mu", c_double), ("kappa", c_double),
    ("IceCompressiveStrength", c_double), ("IceTensileStrength", c_double),
    ("IceShearStrength", c_double),
    ("NACC_beta", c_double), ("NACC_M", c_double), ("NACC_Msq", c_double),
    ("DP_tan_phi", c_double), ("DP_threshold_p", c_double),
    ("cellsize", c_double), ("cellsize_inv", c_double), ("Dp_inv", c_double),
    ("UpdateEveryNthStep", c_int),
    ("IndDiameter", c_double), ("IndRSq", c_double), ("IndVelocity", c_double),
    ("IndDepth", c_double),
    ("xmin", c_double), ("xmax", c_double), ("ymin", c_double), ("ymax", c_double),
    ("nxmin", c_int), ("nxmax", c_int), ("nymin", c_int), ("nymax", c_int),
    ("ParticleVolume", c_double), ("ParticleMass", c_double), ("ParticleViewSize", c_double),
    ("SimulationStep", c_int),
    ("SimulationTime", c_double),
    ("indenter_x", c_double), ("indenter_x_initial", c_double), ("indenter_y", c_double),
    ("indenter_y_initial", c_double),
    ("Volume", c_double),
    ("SetupType", c_int),
    ("GrainVariability", c_double),
    ("UseWarpLevelReduction", c_int),
    ("SortPointsByGridCell", c_int),
]


class SimParams(ctypes.Structure):
    _fields_ = _SIMPARAMS_FIELDS

    def copy(self) -> "SimParams":
        out = SimParams()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(SimParams))
        return out

    def as_bytes(self) -> bytes:
        return ctypes.string_at(ctypes.byref(self), ctypes.sizeof(self))


SIMPARAMS_FIELD_NAMES = [f for f, _ in _SIMPARAMS_FIELDS]

# SoA rows of the point buffer (parameters_sim.h:26-35)
ROW_UTIL, ROW_P, ROW_Q, ROW_JP, ROW_POSX, ROW_VELX, ROW_FE00, ROW_BP00, N_ROWS = (
    0, 1, 2, 3, 4, 6, 8, 12, 16)

_CB = ctypes.CFUNCTYPE(None, c_void_p)


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{os.path.basename(path)} is not built: run `python -m icempm_amd.build` "
            "(the HIP path has no CPU fallback)")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


HIP_EXPORTS = [
    "icempm_create", "icempm_destroy", "icempm_last_error", "icempm_simparams_size",
    "icempm_test", "icempm_synchronize", "icempm_update_constants", "icempm_allocate_arrays",
    "icempm_transfer_points_to_device", "icempm_reset_grid",
    "icempm_reset_indenter_force_accumulator", "icempm_p2g", "icempm_update_nodes",
    "icempm_g2p", "icempm_run_substeps", "icempm_force_history", "icempm_transfer_from_device",
    "icempm_error_code", "icempm_transfer_wait", "icempm_read_error_flag", "icempm_event_record_start",
    "icempm_event_record_stop", "icempm_event_elapsed_ms", "icempm_sort_points",
    "icempm_sort_raw_points", "icempm_sort_raw_points_noalloc", "icempm_set_mode", "icempm_get_mode", "icempm_kernel_timing",
    "icempm_kernel_timing_read", "icempm_read_grid", "icempm_read_points",
    "icempm_write_grid", "icempm_read_bins", "icempm_set_tuning", "icempm_get_tuning",
    "icempm_measure_copy", "icempm_batch_create", "icempm_batch_destroy",
    "icempm_batch_last_error", "icempm_batch_add_case", "icempm_batch_finalize",
    "icempm_batch_run", "icempm_batch_prepare", "icempm_batch_last_ms", "icempm_batch_num_points",
    "icempm_batch_force_history", "icempm_batch_force_bins", "icempm_batch_read_points", "icempm_batch_error_flag",
    "icempm_batch_set_return_map", "icempm_batch_read_harvest", "icempm_batch_map_stats",
    "icempm_batch_map_branch_stats", "icempm_batch_set_map_branches", "icempm_batch_get_map_branches",
    "icempm_batch_autotune_branch_stats",
    "icempm_batch_num_nodes", "icempm_batch_kernel_timing", "icempm_batch_kernel_timing_read",
    "icempm_batch_set_tuning", "icempm_batch_get_tuning", "icempm_batch_resort",
    "icempm_batch_checkpoint", "icempm_batch_restore", "icempm_batch_autotune",
    "icempm_mlp_num_params", "icempm_mlp_create", "icempm_mlp_destroy", "icempm_mlp_last_error",
    "icempm_mlp_get_params", "icempm_mlp_device_params", "icempm_mlp_step", "icempm_mlp_infer",
    "icempm_mfma_probe", "icempm_mfma_probe_k", "icempm_mlp_train", "icempm_device_alloc", "icempm_device_free", "icempm_copy_to_device",
    "icempm_copy_to_host",
]

HOST_EXPORTS = [
    "icempm_host_simparams_size", "icempm_simparams_reset", "icempm_simparams_derive",
    "icempm_simparams_parse_file", "icempm_model_create", "icempm_model_destroy",
    "icempm_model_last_error", "icempm_model_params", "icempm_model_context",
    "icempm_model_load_raw_points", "icempm_model_prepare", "icempm_model_step",
    "icempm_model_wait_cycle", "icempm_model_cycle_time", "icempm_model_host_points",
    "icempm_model_host_force", "icempm_model_set_mode", "icempm_scene_lattice",
]

H5_EXPORTS = [
    "icempm_model_load_raw_points_file", "icempm_model_save_snapshot", "icempm_model_save_pq",
    "icempm_model_read_snapshot",
    "icempm_h5_write_raw_points", "icempm_h5_write_snapshot",
]

_hip = None
_host = None
_h5 = None


def hip():
    """libicempm_hip.so with argument types set."""
    global _hip
    if _hip is not None:
        return _hip
    L = _load(HIP_LIB_PATH)
    P = c_void_p
    _sig(L, "icempm_create", c_int, POINTER(c_void_p))
    _sig(L, "icempm_destroy", None, P)
    _sig(L, "icempm_last_error", c_char_p, P)
    _sig(L, "icempm_simparams_size", c_size_t)
    _sig(L, "icempm_test", c_int, P)
    _sig(L, "icempm_synchronize", c_int, P)
    _sig(L, "icempm_update_constants", c_int, P, c_void_p, c_size_t)
    _sig(L, "icempm_allocate_arrays", c_int, P, c_size_t, c_size_t, c_int, POINTER(c_size_t),
         POINTER(c_size_t), POINTER(POINTER(c_double)), POINTER(POINTER(c_double)))
    _sig(L, "icempm_transfer_points_to_device", c_int, P)
    _sig(L, "icempm_reset_grid", c_int, P)
    _sig(L, "icempm_reset_indenter_force_accumulator", c_int, P)
    _sig(L, "icempm_p2g", c_int, P)
    _sig(L, "icempm_update_nodes", c_int, P, c_double, c_double)
    _sig(L, "icempm_g2p", c_int, P, c_int)
    _sig(L, "icempm_run_substeps", c_int, P, c_int, c_int, POINTER(c_double), c_double, c_double,
         c_int)
    _sig(L, "icempm_force_history", c_int, P, POINTER(c_double), c_int)
    _sig(L, "icempm_transfer_from_device", c_int, P, _CB, c_void_p)
    _sig(L, "icempm_error_code", c_int, P)
    _sig(L, "icempm_transfer_wait", c_int, P)
    _sig(L, "icempm_read_error_flag", c_int, P, POINTER(c_int))
    _sig(L, "icempm_event_record_start", c_int, P)
    _sig(L, "icempm_event_record_stop", c_int, P)
    _sig(L, "icempm_event_elapsed_ms", c_int, P, POINTER(c_float))
    _sig(L, "icempm_sort_points", c_int, P, POINTER(c_int), POINTER(c_int))
    _sig(L, "icempm_sort_raw_points", c_int, P, POINTER(c_float), POINTER(c_int16), c_int,
         c_double, c_int)
    _sig(L, "icempm_sort_raw_points_noalloc", c_int, P, POINTER(c_float), POINTER(c_int16), c_int,
         c_double, c_int)
    _sig(L, "icempm_set_mode", c_int, P, c_int)
    _sig(L, "icempm_get_mode", c_int, P)
    _sig(L, "icempm_kernel_timing", c_int, P, c_int)
    _sig(L, "icempm_kernel_timing_read", c_int, P, POINTER(c_double), POINTER(c_int))
    _sig(L, "icempm_read_grid", c_int, P, POINTER(c_double))
    _sig(L, "icempm_read_points", c_int, P, POINTER(c_double))
    _sig(L, "icempm_write_grid", c_int, P, POINTER(c_double))
    _sig(L, "icempm_read_bins", c_int, P, POINTER(c_double))
    _sig(L, "icempm_set_tuning", c_int, P, c_int, c_int)
    _sig(L, "icempm_get_tuning", c_int, P, c_int)
    _sig(L, "icempm_measure_copy", c_int, P, c_size_t, c_int, POINTER(c_double))
    _sig(L, "icempm_batch_create", c_int, POINTER(c_void_p), c_int)
    _sig(L, "icempm_batch_destroy", None, P)
    _sig(L, "icempm_batch_last_error", c_char_p, P)
    _sig(L, "icempm_batch_add_case", c_int, P, c_void_p, c_size_t, POINTER(c_double), c_size_t)
    _sig(L, "icempm_batch_finalize", c_int, P)
    _sig(L, "icempm_batch_run", c_int, P, c_int, c_int)
    _sig(L, "icempm_batch_prepare", c_int, P, c_int)
    _sig(L, "icempm_batch_last_ms", c_int, P, POINTER(c_float))
    _sig(L, "icempm_batch_num_points", c_int, P)
    _sig(L, "icempm_batch_force_history", c_int, P, c_int, POINTER(c_double), c_int)
    _sig(L, "icempm_batch_force_bins", c_int, P, c_int, POINTER(c_double), c_int)
    _sig(L, "icempm_batch_read_points", c_int, P, c_int, POINTER(c_double), c_size_t)
    _sig(L, "icempm_batch_error_flag", c_int, P, POINTER(c_int))
    _sig(L, "icempm_batch_set_return_map", c_int, P, c_int, POINTER(c_float), POINTER(c_float),
         c_double, c_int)
    _sig(L, "icempm_batch_read_harvest", c_int, P, POINTER(c_float), POINTER(c_float), c_int,
         POINTER(c_uint64))
    _sig(L, "icempm_batch_map_stats", c_int, P, POINTER(c_uint64), c_int)
    _sig(L, "icempm_batch_map_branch_stats", c_int, P, POINTER(c_uint64), c_int)
    _sig(L, "icempm_batch_set_map_branches", c_int, P, c_int)
    _sig(L, "icempm_batch_get_map_branches", c_int, P)
    _sig(L, "icempm_batch_autotune_branch_stats", c_int, P, POINTER(c_uint64))
    _sig(L, "icempm_batch_num_nodes", c_long, P)
    _sig(L, "icempm_batch_kernel_timing", c_int, P, c_int)
    _sig(L, "icempm_batch_kernel_timing_read", c_int, P, POINTER(c_double), POINTER(c_int))
    _sig(L, "icempm_batch_set_tuning", c_int, P, c_int, c_int)
    _sig(L, "icempm_batch_get_tuning", c_int, P, c_int)
    _sig(L, "icempm_batch_resort", c_int, P)
    _sig(L, "icempm_batch_checkpoint", c_int, P)
    _sig(L, "icempm_batch_restore", c_int, P)
    _sig(L, "icempm_batch_autotune", c_int, P, c_int, POINTER(c_double), c_int, POINTER(c_int))
    fp = POINTER(c_float)
    _sig(L, "icempm_mlp_num_params", c_int)
    _sig(L, "icempm_mlp_create", c_int, POINTER(c_void_p), fp)
    _sig(L, "icempm_mlp_destroy", None, P)
    _sig(L, "icempm_mlp_last_error", c_char_p, P)
    _sig(L, "icempm_mlp_get_params", c_int, P, fp)
    _sig(L, "icempm_mlp_device_params", c_void_p, P)
    _sig(L, "icempm_mlp_step", c_int, P, c_void_p, c_void_p, c_int, c_int, c_float, fp, fp)
    _sig(L, "icempm_mlp_infer", c_int, P, c_void_p, c_int, c_void_p)
    _sig(L, "icempm_mfma_probe", c_int, fp, fp, fp)
    _sig(L, "icempm_mfma_probe_k", c_int, fp, fp, fp, c_int)
    _sig(L, "icempm_mlp_train", c_int, fp, fp, c_int, c_int, c_int, c_float, c_ulonglong, fp, fp, fp, c_int, fp)
    _sig(L, "icempm_device_alloc", c_int, c_size_t, POINTER(c_void_p))
    _sig(L, "icempm_device_free", None, c_void_p)
    _sig(L, "icempm_copy_to_device", c_int, c_void_p, c_void_p, c_size_t)
    _sig(L, "icempm_copy_to_host", c_int, c_void_p, c_void_p, c_size_t)
    _hip = L
    return L


def host():
    """libicempm_host.so with argument types set."""
    global _host
    if _host is not None:
        return _host
    hip()
    L = _load(HOST_LIB_PATH)
    P = c_void_p
    _sig(L, "icempm_host_simparams_size", c_size_t)
    _sig(L, "icempm_simparams_reset", None, c_void_p)
    _sig(L, "icempm_simparams_derive", None, c_void_p)
    _sig(L, "icempm_simparams_parse_file", c_int, c_void_p, c_char_p, c_char_p, c_size_t,
         c_char_p, c_size_t)
    _sig(L, "icempm_model_create", c_int, POINTER(c_void_p))
    _sig(L, "icempm_model_destroy", None, P)
    _sig(L, "icempm_model_last_error", c_char_p, P)
    _sig(L, "icempm_model_params", c_void_p, P)
    _sig(L, "icempm_model_context", c_void_p, P)
    _sig(L, "icempm_model_load_raw_points", c_int, P, POINTER(c_float), POINTER(c_int16), c_int,
         c_float)
    _sig(L, "icempm_model_prepare", c_int, P)
    _sig(L, "icempm_model_step", c_int, P, POINTER(c_int))
    _sig(L, "icempm_model_wait_cycle", c_int, P)
    _sig(L, "icempm_model_cycle_time", c_float, P)
    _sig(L, "icempm_model_host_points", POINTER(c_double), P)
    _sig(L, "icempm_model_host_force", POINTER(c_double), P)
    _sig(L, "icempm_model_set_mode", c_int, P, c_int)
    _sig(L, "icempm_scene_lattice", c_long, c_int, c_int, c_int, c_double, c_double, c_uint64,
         POINTER(c_float), POINTER(c_int16))
    _host = L
    return L


def h5_built() -> bool:
    return os.path.exists(H5_LIB_PATH)


def h5():
    """libicempm_h5.so (SnapshotManager's HDF5 files) with argument types set."""
    global _h5
    if _h5 is not None:
        return _h5
    host()
    L = _load(H5_LIB_PATH)
    P = c_void_p
    _sig(L, "icempm_model_load_raw_points_file", c_int, P, c_char_p)
    _sig(L, "icempm_model_save_snapshot", c_int, P, c_char_p)
    _sig(L, "icempm_model_save_pq", c_int, P, c_char_p)
    _sig(L, "icempm_model_read_snapshot", c_int, P, c_char_p)
    _sig(L, "icempm_h5_write_raw_points", c_int, c_char_p, POINTER(c_float), POINTER(c_int16), c_int,
         c_float, c_int)
    _sig(L, "icempm_h5_write_snapshot", c_int, c_char_p, c_void_p, c_size_t, POINTER(c_double),
         POINTER(c_double))
    _h5 = L
    return L


def check_ctx(rc: int, ctx) -> None:
    if rc != ICEMPM_OK:
        msg = hip().icempm_last_error(ctx)
        raise IcempmError(rc, (msg or b"").decode(errors="replace"))
