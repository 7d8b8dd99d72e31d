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
