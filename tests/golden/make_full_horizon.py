"""Golden vectors of the full-horizon runs, written by the CPU oracle.

  indent60k_20000.npz    the 60k-point case (ice block 2.4 m x 0.9 m, grid
                         256 x 128, dt 2e-6 s) over its full 20,000 substeps
                         = 8 Model::Step cycles of 2,500 (model.cpp:11-47),
                         by the sequential oracle: the bit pattern an ORDERED
                         run must reproduce.
  indent1.6M_500.npz     the 1.6M-point case (5 m x 2 m, grid 1200 x 500,
                         dt 1e-6 s) over its first 500 substeps, by the
                         OpenMP oracle (8 threads; its P2G sums per-thread
                         bands, so it differs from the sequential oracle
                         only in the order of those sums): the reference a
                         FAST run is held to within tolerance.

Each file holds the SHA-256 of the ingested point SoA (so a test can show
its own ingest starts from the same bits), the per-substep total indenter
force, the Indenter_2D bins of every output cycle, the number of crushed
points after every substep and the indices of the crushed points every
`checkpoint` substeps, the SHA-256 of the final SoA, and the final positions
of every `stride`-th point.

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_full_horizon.py [indent-60k|indent-1.6M|all]
The 60k file takes about 4 minutes, the 1.6M file about 3 (8 cores).
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from icempm_amd import INDENT_1M6, INDENT_60K, SimParams, configure_params, lattice_points  # noqa: E402

SPECS = {
    "indent-60k": dict(case=INDENT_60K, nsub=20000, threads=1, checkpoint=500, stride=7,
                       out="indent60k_20000.npz"),
    "indent-1.6M": dict(case=INDENT_1M6, nsub=500, threads=8, checkpoint=100, stride=97,
                        out="indent1.6M_500.npz"),
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pitch_of(n: int) -> int:
    return max(64, (n + 63) // 64 * 64)


def make(name: str) -> str:
    s = SPECS[name]
    case, nsub, threads = s["case"], s["nsub"], s["threads"]
    p = SimParams()
    configure_params(p, case)
    xy, g = lattice_points(case)
    pts = oracle.ingest(p, xy, g, case.volume, pitch=pitch_of(case.n_points))
    p.nGridPitch = pitch_of(p.GridTotal)
    n = case.n_points
    N = p.UpdateEveryNthStep
    init = sha(pts)
    grid = oracle.new_grid(p)
    bins = np.zeros(2 * p.n_indenter_subdivisions)
    hist = np.zeros((nsub, 2))
    counts = np.zeros(nsub, dtype=np.int32)
    cyc_bins, cr_idx, cr_off = [], [], [0]
    t0 = time.time()
    for i in range(nsub):
        h, err = oracle.run(p, pts, grid, bins, 1, nthreads=threads)
        if err:
            raise RuntimeError(f"{name}: out-of-grid point at substep {i}")
        hist[i] = h[0]
        crushed = pts[0].view(np.uint8)[:n]
        counts[i] = int(crushed.sum())
        if (i + 1) % N == 0 or i + 1 == nsub:
            cyc_bins.append(bins.copy())
        if (i + 1) % s["checkpoint"] == 0:
            idx = np.flatnonzero(crushed).astype(np.int32)
            cr_idx.append(idx)
            cr_off.append(cr_off[-1] + idx.size)
        if (i + 1) % 1000 == 0:
            print(f"{name}: substep {i + 1}, crushed {counts[i]}, {time.time() - t0:.0f} s", flush=True)
    out = os.path.join(HERE, s["out"])
    np.savez_compressed(
        out, init_sha256=np.array(init), final_sha256=np.array(sha(pts)), hist=hist,
        bins=np.array(cyc_bins), crushed_count=counts,
        crushed_idx=np.concatenate(cr_idx) if cr_idx else np.zeros(0, np.int32),
        crushed_off=np.array(cr_off, dtype=np.int64), checkpoint=np.array(s["checkpoint"]),
        stride=np.array(s["stride"]), pos_final=pts[4:6, :n:s["stride"]].copy(),
        threads=np.array(threads), simulation_time=np.array(p.SimulationTime))
    print(f"{name}: wrote {out} ({os.path.getsize(out)} bytes), first crush at substep "
          f"{int(np.argmax(counts > 0)) + 1 if counts.any() else None}", flush=True)
    return out


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    for nm in (SPECS if which == "all" else [which]):
        make(nm)
