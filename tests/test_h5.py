"""HDF5 raw point files in the reference's input format (no GPU): the
writer's file layout as h5dump sees it, its data read back by h5dump, and
the same structure as the reference's own input geometry files."""
import os

import numpy as np
import pytest
from h5util import H5DUMP, attr_values, binary, header, structure

from icempm_amd import INDENT_TINY, _lib, lattice_points
from icempm_amd.model import write_raw_points

REF_INPUT = "/root/reference/geometries_2d/2d_70k.h5"

pytestmark = pytest.mark.skipif(not (_lib.h5_built() and os.path.exists(H5DUMP)),
                                reason="libicempm_h5.so or h5dump not available")


@pytest.mark.parametrize("deflate", [0, 7])
def test_raw_points_file_layout_and_data(tmp_path, deflate):
    xy, _ = lattice_points(INDENT_TINY)
    grains = np.random.default_rng(1).integers(-32768, 32767, size=xy.shape[0]).astype(np.int16)
    path = tmp_path / "raw.h5"
    write_raw_points(str(path), xy, grains, 2.5, deflate)
    s = structure(path)
    assert ("dataset", "Points_Raw_2D", "H5T_IEEE_F32LE", 2) in s
    assert ("dataset", "GrainIDs", "H5T_STD_I16LE", 1) in s
    assert ("attribute", "volume", "H5T_IEEE_F32LE") in s
    text = header(path)
    assert f"( {xy.shape[0]}, 2 )" in text and f"( {xy.shape[0]} )" in text
    assert np.array_equal(binary(path, "/Points_Raw_2D", "<f4", tmp_path).reshape(-1, 2), xy)
    assert np.array_equal(binary(path, "/GrainIDs", "<i2", tmp_path), grains)
    assert attr_values(path, "/GrainIDs/volume") == [2.5]


@pytest.mark.skipif(not os.path.exists(REF_INPUT), reason="reference input geometry not present")
def test_same_structure_as_reference_input(tmp_path):
    xy, g = lattice_points(INDENT_TINY)
    path = tmp_path / "raw.h5"
    write_raw_points(str(path), xy, g, INDENT_TINY.volume, 7)
    assert structure(path) == structure(REF_INPUT)


def test_snapshot_from_host_data(tmp_path):
    """icempm_h5_write_snapshot (cm's batch mode): the single-case snapshot
    layout -- Indenter_2D, Params with its attributes, Points -- written from
    caller-held arrays, named for the animation frame, read back unchanged;
    bad arguments are refused before anything is written."""
    import ctypes

    from icempm_amd import SimParams, configure_params

    p = SimParams()
    configure_params(p, INDENT_TINY)
    p.nPtsPitch = 128
    p.UpdateEveryNthStep = 10
    p.SimulationStep = 30
    p.SimulationTime = 30 * p.InitialTimeStep
    p.indenter_x = p.indenter_x_initial + p.SimulationTime * p.IndVelocity
    rng = np.random.default_rng(5)
    bins = rng.standard_normal(2 * p.n_indenter_subdivisions)
    pts = rng.standard_normal(16 * p.nPtsPitch)
    L = _lib.h5()
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.icempm_h5_write_snapshot(os.fsencode(str(tmp_path)), ctypes.byref(p), ctypes.sizeof(p),
                                    bins.ctypes.data_as(dp), pts.ctypes.data_as(dp))
    assert rc == 0
    path = tmp_path / "snapshots" / "d00003.h5"
    s = structure(path)
    for name, typ in (("Indenter_2D", "H5T_IEEE_F64LE"), ("Points", "H5T_IEEE_F64LE")):
        assert ("dataset", name, typ, 1) in s
    assert ("attribute", "SimulationTime", "H5T_IEEE_F64LE") in s
    assert ("attribute", "nPtsPitch", "H5T_STD_I64LE") in s
    assert np.array_equal(binary(path, "/Indenter_2D", "<f8", tmp_path), bins)
    assert np.array_equal(binary(path, "/Points", "<f8", tmp_path), pts)
    back = SimParams.from_buffer_copy(binary(path, "/Params", "u1", tmp_path).tobytes())
    assert bytes(back) == bytes(p)
    for a in ("indenter_x", "SimulationTime"):   # h5dump prints them to ~15 digits
        assert abs(attr_values(path, "/Params/" + a)[0] - getattr(p, a)) <= 1e-12 * abs(getattr(p, a)), a
    # refused: wrong struct size, no frame step, null arrays
    assert L.icempm_h5_write_snapshot(os.fsencode(str(tmp_path)), ctypes.byref(p), ctypes.sizeof(p) - 8,
                                      bins.ctypes.data_as(dp), pts.ctypes.data_as(dp)) != 0
    q = SimParams.from_buffer_copy(bytes(p))
    q.UpdateEveryNthStep = 0
    assert L.icempm_h5_write_snapshot(os.fsencode(str(tmp_path / "x")), ctypes.byref(q), ctypes.sizeof(q),
                                      bins.ctypes.data_as(dp), pts.ctypes.data_as(dp)) != 0
    assert L.icempm_h5_write_snapshot(os.fsencode(str(tmp_path / "x")), ctypes.byref(p), ctypes.sizeof(p),
                                      None, pts.ctypes.data_as(dp)) != 0
    assert not (tmp_path / "x").exists()
