"""The C++ drop-in surface end to end: a frontend program written like the
reference's CLI (JSON config -> LoadRawPoints -> Model::Step loop with the
readback callback) runs on the GPU; its last readback must equal the oracle
bit for bit in ORDERED mode."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from icempm_amd import INDENT_TINY, SimParams, configure_params, lattice_points

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "drop_in_driver")


def case_json(case, path):
    cfg = {
        "InputRawPoints": "lattice",
        "InitialTimeStep": case.dt, "YoungsModulus": case.youngs, "GridX": case.grid_x,
        "GridY": case.grid_y, "GridXDimension": case.grid_x * case.h, "SimulationEndTime": 1e9,
        "PoissonsRatio": case.poisson, "Gravity": case.gravity, "Density": case.density,
        "IndDiameter": case.ind_diameter, "IndVelocity": case.ind_velocity,
        "IndDepth": case.ind_depth, "IceCompressiveStrength": case.compressive,
        "IceTensileStrength": case.tensile, "IceShearStrength": case.shear,
        "DP_phi": case.dp_phi_deg, "DP_threshold_p": case.dp_threshold,
        "GrainVariability": case.grain_variability,
    }
    # floats are written with repr() so that strtod reads back the same doubles
    text = "{" + ", ".join(f'"{k}": ' + (repr(v) if isinstance(v, float) else json.dumps(v))
                           for k, v in cfg.items()) + "}"
    path.write_text(text)


def test_cli_like_frontend_ordered(gpu_available, tmp_path):
    assert os.path.exists(DRIVER), "build the driver first (__graft_entry__.build)"
    case = INDENT_TINY
    cfg = tmp_path / "case.json"
    case_json(case, cfg)
    out = tmp_path / "out.bin"
    print("drop-in driver", case.name, flush=True)
    r = subprocess.run([DRIVER, str(cfg), str(case.ncx), str(case.ncy), str(case.per_side),
                        repr(case.h), str(case.seed), "1", "1", str(out)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout

    p = SimParams()
    configure_params(p, case)
    xy, g = lattice_points(case)
    pitch = ((case.n_points + 63) // 64) * 64
    pts = oracle.ingest(p, xy, g, case.volume, pitch)
    p.nGridPitch = ((p.GridTotal + 63) // 64) * 64
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    N = p.UpdateEveryNthStep
    oracle.run(p, pts, grid, bins, N)

    raw = np.fromfile(out, dtype=np.float64)
    npts = 16 * pitch
    assert raw.size == npts + bins.size + 2
    assert np.array_equal(raw[:npts].view(np.uint64), pts.ravel().view(np.uint64))
    assert np.array_equal(raw[npts:npts + bins.size], bins)
    assert raw[-2] == N and raw[-1] == p.SimulationTime
