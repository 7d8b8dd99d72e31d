"""SnapshotManager's HDF5 files on the GPU path: LoadRawPoints from a raw
point file equals the array ingest, and SaveSnapshot / SavePQ write the
reference's layout (datasets, types, attributes) with the readback buffers'
contents, as h5dump reads them back."""
import ctypes
import os

import numpy as np
import pytest
from h5util import H5DUMP, attr_values, binary, structure

from icempm_amd import INDENT_TINY, MODE_ORDERED, Model, _lib, configure_params, lattice_points
from icempm_amd.model import soa_views, write_raw_points

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (_lib.h5_built() and os.path.exists(H5DUMP)),
                                 reason="libicempm_h5.so or h5dump not available")]


def test_load_raw_points_file_equals_arrays(gpu_available, tmp_path):
    print("snapshot: load raw file", flush=True)
    case = INDENT_TINY
    xy, g = lattice_points(case)
    g = (np.arange(xy.shape[0]) % 1000).astype(np.int16)
    path = tmp_path / "raw.h5"
    write_raw_points(str(path), xy, g, case.volume, 7)
    a, b = Model(MODE_ORDERED), Model(MODE_ORDERED)
    try:
        configure_params(a.prms, case)
        configure_params(b.prms, case)
        a.load_raw_points_file(str(path))
        b.load_raw_points(xy, g, case.volume)
        assert bytes(a.prms) == bytes(b.prms)
        assert np.array_equal(a.points(), b.points())
    finally:
        a.close()
        b.close()


def test_snapshot_and_pq_files(gpu_available, tmp_path):
    print("snapshot: save", flush=True)
    case = INDENT_TINY
    m = Model(MODE_ORDERED)
    try:
        m.load_case(case)
        m.step()                  # one output cycle, then the readback
        m.wait_cycle()
        p = m.prms
        path = m.save_snapshot(str(tmp_path))
        assert os.path.exists(path) and path.endswith("d00001.h5")
        s = structure(path)
        for entry in [("dataset", "Indenter_2D", "H5T_IEEE_F64LE", 1), ("dataset", "Params", "H5T_STD_B8LE", 1),
                      ("dataset", "Points", "H5T_IEEE_F64LE", 1)]:
            assert entry in s, entry
        names = {e[1] for e in s if e[0] == "attribute"}
        assert names == {"indenter_x", "indenter_y", "SimulationTime", "GridX", "GridY", "nPts", "nPtsPitch",
                         "UpdateEveryNthStep", "n_indenter_subdivisions", "cellsize", "IndDiameter",
                         "InitialTimeStep", "SetupType", "Volume"}
        assert attr_values(path, "/Params/nPts") == [p.nPts]
        assert attr_values(path, "/Params/nPtsPitch") == [p.nPtsPitch]
        assert attr_values(path, "/Params/SimulationTime") == [pytest.approx(p.SimulationTime, rel=1e-15)]
        pts = binary(path, "/Points", "<f8", tmp_path)
        host = m.host_points()
        assert np.array_equal(pts.view(np.uint64), host.ravel().view(np.uint64))
        assert np.array_equal(binary(path, "/Indenter_2D", "<f8", tmp_path), m.host_force())
        prm = binary(path, "/Params", "u1", tmp_path)
        assert prm.tobytes() == bytes(p)

        pq = m.save_pq(str(tmp_path))
        v = soa_views(host, p.nPts)
        cr = v["crushed"] != 0
        both = np.stack([v["P"], v["Q"]], axis=1)
        intact = binary(pq, "/PQ_intact", "<f8", tmp_path).reshape(-1, 2)
        crushed = binary(pq, "/PQ_crushed", "<f8", tmp_path).reshape(-1, 2) if cr.any() else np.zeros((0, 2))
        assert np.array_equal(intact, both[~cr])
        assert np.array_equal(crushed, both[cr])
        assert ("attribute", "nPts", "H5T_STD_I32LE") in structure(pq)
    finally:
        m.close()


def test_resume_from_snapshot_continues_bit_exact(gpu_available, tmp_path):
    """ORDERED: two cycles straight equal one cycle, a snapshot, a new model
    resumed from it and one more cycle."""
    print("snapshot: resume", flush=True)
    case = INDENT_TINY
    a = Model(MODE_ORDERED)
    b = Model(MODE_ORDERED)
    try:
        a.load_case(case)
        a.step()
        a.wait_cycle()
        path = a.save_snapshot(str(tmp_path))
        a.step()
        a.wait_cycle()
        configure_params(b.prms, case)
        b.read_snapshot(path)
        assert b.prms.SimulationStep == a.prms.UpdateEveryNthStep
        b.step()
        b.wait_cycle()
        assert b.prms.SimulationStep == a.prms.SimulationStep
        assert b.prms.SimulationTime == a.prms.SimulationTime
        assert np.array_equal(b.host_points().view(np.uint64), a.host_points().view(np.uint64))
        assert np.array_equal(b.host_force(), a.host_force())
    finally:
        a.close()
        b.close()
