"""The cm frontend (cli/main.cpp's behaviour): a JSON configuration and an
HDF5 raw point file in, one HDF5 snapshot per output cycle out; in ORDERED
mode the snapshot equals the CPU oracle's state bit for bit."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest
from h5util import H5DUMP, binary

import oracle
from icempm_amd import INDENT_TINY, SimParams, _lib, configure_params, lattice_points
from icempm_amd.build import CLI_BIN
from icempm_amd.model import write_raw_points

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(CLI_BIN) and os.path.exists(H5DUMP)),
                                 reason="cm or h5dump not built / available")]


def test_cm_writes_oracle_snapshots(gpu_available, tmp_path):
    print("cli: cm, 2 cycles ORDERED", flush=True)
    case = INDENT_TINY
    xy, g = lattice_points(case)
    raw = tmp_path / "raw.h5"
    write_raw_points(str(raw), xy, g, case.volume, 7)
    p = SimParams()
    configure_params(p, case)
    cfg = {
        "InitialTimeStep": case.dt, "YoungsModulus": case.youngs, "GridX": case.grid_x,
        "GridY": case.grid_y, "GridXDimension": p.GridXDimension, "SimulationEndTime": 1.0,
        "PoissonsRatio": case.poisson, "Gravity": case.gravity, "Density": case.density,
        "IndDiameter": case.ind_diameter, "IndVelocity": case.ind_velocity, "IndDepth": case.ind_depth,
        "IceCompressiveStrength": case.compressive, "IceTensileStrength": case.tensile,
        "IceShearStrength": case.shear, "DP_phi": case.dp_phi_deg, "DP_threshold_p": case.dp_threshold,
        "GrainVariability": case.grain_variability, "InputRawPoints": str(raw),
    }
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    out = tmp_path / "snap"
    r = subprocess.run([CLI_BIN, str(tmp_path / "c.json"), "--snapshots", str(out), "--ordered", "--pq",
                        "--cycles", "2"], stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=180)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "cm done" in r.stdout
    for f in ("snapshots/d00001.h5", "snapshots/d00002.h5", "pq/pq00001.h5", "pq/pq00002.h5"):
        assert (out / f).exists(), f

    # the oracle: the same ingest and 2 cycles of Model::Step
    snap = out / "snapshots" / "d00002.h5"
    prm = SimParams.from_buffer_copy(binary(snap, "/Params", "u1", tmp_path).tobytes())
    pts = oracle.ingest(p, xy, g, case.volume, prm.nPtsPitch)
    p.nGridPitch = prm.nGridPitch
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    N = p.UpdateEveryNthStep
    _, err = oracle.run(p, pts, grid, bins, 2 * N)
    assert err == 0
    assert prm.SimulationStep == p.SimulationStep and prm.SimulationTime == p.SimulationTime
    dev = binary(snap, "/Points", "<f8", tmp_path).reshape(16, -1)
    assert np.array_equal(dev.view(np.uint64), pts.view(np.uint64))
    assert np.array_equal(binary(snap, "/Indenter_2D", "<f8", tmp_path), bins)
