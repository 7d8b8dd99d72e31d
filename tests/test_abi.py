"""The C ABI libraries load and export every entry point their headers
declare; the SimParams layout agrees between the C++ struct, the ctypes
mirror and the oracle.  No GPU calls."""
import ctypes
import os
import re

import pytest

import oracle
from icempm_amd import H5_EXPORTS, HIP_EXPORTS, HOST_EXPORTS, SimParams, _lib, hip, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(icempm_\w+)\s*\(", text)
    # drop typedef'd function-pointer names and macros
    return sorted(set(n for n in names if not n.startswith("icempm_callback")))


def h5_loader():
    if not _lib.h5_built():
        pytest.skip("libicempm_h5.so not built (no HDF5 C library)")
    return _lib.h5()


@pytest.mark.parametrize("header,loader,listed", [
    ("icempm.h", hip, HIP_EXPORTS),
    ("icempm_host.h", host, HOST_EXPORTS),
    ("icempm_h5.h", h5_loader, H5_EXPORTS),
])
def test_exports(header, loader, listed):
    names = declared(header)
    assert names, header
    L = loader()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(listed) == names


def test_simparams_layout_agrees():
    assert ctypes.sizeof(SimParams) == 432
    assert hip().icempm_simparams_size() == 432
    assert host().icempm_host_simparams_size() == 432
    assert oracle.lib().oracle_simparams_size() == 432
    offs = (ctypes.c_long * 128)()
    n = oracle.lib().oracle_simparams_offsets(offs, 128)
    ours = [getattr(SimParams, f).offset for f, _ in SimParams._fields_]
    assert n == len(ours)
    assert list(offs[:n]) == ours


def test_reset_defaults():
    p = SimParams()
    host().icempm_simparams_reset(ctypes.byref(p))
    assert p.n_indenter_subdivisions == 100
    assert p.GridX == 128 and p.GridY == 55 and p.GridTotal == 128 * 55
    assert p.InitialTimeStep == 3e-5 and p.UpdateEveryNthStep == 166
    assert p.tpb_P2G == 256 and p.SortPointsByGridCell == 1
    assert p.grid_array is None and p.pts_array is None


def test_parse_file(tmp_path):
    cfg = tmp_path / "c.json"
    cfg.write_text('{"InputRawPoints":"pts.h5", "InitialTimeStep":2.0e-6, "GridX":256, '
                   '"GridY":128, "GridXDimension":3.072, "DP_phi":70, "tpb_P2G":64, '
                   '"Extra":{"a":[1,2]}, "IndDepth":0.1}')
    p = SimParams()
    host().icempm_simparams_reset(ctypes.byref(p))
    raw = ctypes.create_string_buffer(256)
    err = ctypes.create_string_buffer(256)
    rc = host().icempm_simparams_parse_file(ctypes.byref(p), str(cfg).encode(), raw, 256, err, 256)
    assert rc == 0, err.value
    assert raw.value == b"pts.h5"
    assert p.UpdateEveryNthStep == 2500 and p.GridTotal == 256 * 128
    assert p.cellsize == pytest.approx(0.012, rel=1e-15)
    assert p.DP_tan_phi == pytest.approx(2.7474774194546216, rel=1e-14)
    assert p.tpb_P2G == 64 and p.IndDepth == 0.1


@pytest.mark.parametrize("text,msg", [
    (None, b"configuration file is not found"),
    ("not json", b"configuration file is not JSON"),
    ('{"GridX": 10}', b"config parameter missing"),
])
def test_parse_file_errors(tmp_path, text, msg):
    path = tmp_path / "c.json"
    if text is not None:
        path.write_text(text)
    p = SimParams()
    host().icempm_simparams_reset(ctypes.byref(p))
    err = ctypes.create_string_buffer(256)
    rc = host().icempm_simparams_parse_file(ctypes.byref(p), str(path).encode(), None, 0, err, 256)
    assert rc != 0 and err.value == msg


def test_scene_lattice_is_seeded():
    from icempm_amd import INDENT_TINY, lattice_points
    xy1, g1 = lattice_points(INDENT_TINY)
    xy2, _ = lattice_points(INDENT_TINY)
    assert xy1.shape == (INDENT_TINY.n_points, 2) and (g1 == 0).all()
    assert (xy1 == xy2).all()
    h = INDENT_TINY.h
    # each point stays within a quarter cell of its lattice site
    k = INDENT_TINY.per_side
    i = np.arange(INDENT_TINY.n_points)
    a = i % k
    b = (i // k) % k
    cx = (i // (k * k)) % INDENT_TINY.ncx
    cy = i // (k * k * INDENT_TINY.ncx)
    sx = (cx + (a + 0.5) / k) * h
    sy = (cy + (b + 0.5) / k) * h
    assert np.abs(xy1[:, 0] - sx).max() <= 0.25 * h * (1 + 1e-6)
    assert np.abs(xy1[:, 1] - sy).max() <= 0.25 * h * (1 + 1e-6)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "HIP_LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        _lib._load(_lib.HIP_LIB_PATH)


import numpy as np  # noqa: E402


def test_host_guard_turns_exceptions_into_status(tmp_path):
    """No C++ exception leaves a C entry point: the guard every entry point of
    libicempm_hip.so runs its body in returns ICEMPM_E_HOST with the text."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "host_guard_check")
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", f"-I{os.path.join(ROOT, 'icempm_amd', 'csrc')}",
                    os.path.join(ROOT, "tests", "cpp", "host_guard_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 6, r.stdout
    assert _lib.ICEMPM_E_HOST == 6


def test_every_hip_entry_point_is_guarded():
    """Each C entry point of libicempm_hip.so with a body of its own runs it
    behind host_guard (one-line accessors excepted): a scan of the sources."""
    unguarded = []
    for src in ("api.hip", "batch.hip"):
        lines = open(os.path.join(ROOT, "icempm_amd", "csrc", src)).read().split("\n")
        in_c = False
        for i, line in enumerate(lines):
            if line.startswith('extern "C" {'):
                in_c = True
            elif line.startswith('}  // extern "C"'):
                in_c = False
            m = re.match(r"^int (icempm_\w+)\(", line)
            if not (in_c and m) or line.rstrip().endswith("}"):
                continue
            j = i
            while not lines[j].rstrip().endswith(") {"):
                j += 1
            first_param = "".join(lines[i:j + 1]).split("(", 1)[1].split(",")[0]
            if first_param.startswith("const icempm_"):
                continue   # read-only accessors
            if not lines[j + 1].strip().startswith("return host_guard("):
                unguarded.append(f"{src}:{i + 1} {m.group(1)}")
    assert not unguarded, unguarded
