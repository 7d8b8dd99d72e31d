// internal.h -- launchers shared between the translation units of
// libicempm_hip.so.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "host_guard.h"
#include "mpm_device.h"

namespace icempm {

// Geometry of the shared-memory window that one P2G / G2P workgroup uses:
// a chunk of at most kChunk consecutive particles whose base cells lie in
// one cell row at build time; the window spans kWinX x kWinY grid nodes from
// the chunk's minimum base cell.
constexpr int kChunk = 256;
constexpr int kWinX = 128;
constexpr int kWinY = 4;

// ---- kernels.hip ------------------------------------------------------------
hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s);
hipError_t launch_fill_i32(int *p, size_t n, int v, hipStream_t s);

// ---- particles_exact.hip / particles_fast.hip ---------------------------------
// The particle pass.  do_g2p: gather from gridV (velocities in rows 1-2) and
// update the particles; do_p2g: scatter the (updated) particles into gridOut
// (mass, momentum).  gridV and gridOut may be the same buffer only when one
// of the two passes is off.  exact:: is built without FMA contraction
// (ORDERED mode), fast:: with it (FAST mode).  kind selects the kernel
// variant: 0 workgroup windows (shared-memory gather and scatter), 1 wavefront
// windows (direct gather, per-wavefront scatter), 2 as 1 with the stencil
// order rotated per lane.
namespace exact {
hipError_t launch_particles(const KParams &k, bool do_g2p, bool do_p2g, const double *gridV,
                            double *gridOut, const int *chunkStart, int nChunks, int recordPQ,
                            hipStream_t s, int kind);
}
namespace fast {
hipError_t launch_particles(const KParams &k, bool do_g2p, bool do_p2g, const double *gridV,
                            double *gridOut, const int *chunkStart, int nChunks, int recordPQ,
                            hipStream_t s, int kind);
}

// Grid update.  ordered: write each node's contact force and bin to
// nodeBin/nodeF for a later ordered fold; otherwise accumulate into k.bins
// and hist2 (if not null) with atomics.  zeroGrid (may be null): a second
// grid whose 3 rows are cleared in the same pass.
hipError_t launch_update(const KParams &k, double *grid, double ix, double iy, double *hist2,
                         double *zeroGrid, bool ordered, int *nodeBin, double *nodeF,
                         hipStream_t s);

// ORDERED-mode P2G: contributions in particle order per node.
struct OrderedScratch {
    double *contrib = nullptr;   // 27 doubles per particle
    int *cnode = nullptr;        // 9 node ids per particle (-1: none)
    int *count = nullptr;        // gridTotal + 1
    int *start = nullptr;        // gridTotal + 1
    int *cursor = nullptr;       // gridTotal
    int *list = nullptr;         // 9 per particle
    int *nodeBin = nullptr;      // gridTotal
    double *nodeF = nullptr;     // 2 per node
    int *flags = nullptr;        // gridTotal + 1
    int *cpos = nullptr;         // gridTotal + 1
    int *contacts = nullptr;     // gridTotal
    int *scanTmp = nullptr;      // scan scratch
    size_t nPts = 0, nGrid = 0;
};
hipError_t ordered_p2g(const KParams &k, double *grid, OrderedScratch &o, hipStream_t s);
hipError_t ordered_fold_forces(const KParams &k, OrderedScratch &o, double *hist2,
                               hipStream_t s);

hipError_t launch_hello(int *flag, hipStream_t s);
hipError_t launch_copy(const double *src, double *dst, size_t n_doubles, hipStream_t s);

// ---- sort.hip -----------------------------------------------------------------
// Exclusive scan of n ints (in and out may alias).  tmp must hold
// scan_tmp_ints(n) ints.
size_t scan_tmp_ints(size_t n);
hipError_t exclusive_scan(const int *in, int *out, size_t n, int *tmp, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs by signed 32-bit key.
// Buffers: keys/vals (in, out), keys2/vals2 (scratch of n), tmp of
// radix_tmp_ints(n) ints.  The result ends in keys/vals.
size_t radix_tmp_ints(size_t n);
hipError_t radix_sort_pairs(int *keys, int *vals, int *keys2, int *vals2, size_t n, int *tmp,
                            hipStream_t s);

// key_i of the particles in a point SoA
hipError_t launch_point_keys(const double *pts, size_t pp, int n, double hinv, int gridX,
                             int *keys, int *vals, hipStream_t s);
// key_i of raw float (x, y) pairs
hipError_t launch_raw_keys(const float *xy, int n, double hinv, int gridX, int *keys, int *vals,
                           hipStream_t s);
// dst = src permuted: dst slot k <- src slot perm[k], for all 16 rows
hipError_t launch_permute_points(const double *src, double *dst, size_t pp, int n,
                                 const int *perm, hipStream_t s);
hipError_t launch_permute_raw(const float *xy, const int16_t *g, float *xy2, int16_t *g2, int n,
                              const int *perm, hipStream_t s);

}  // namespace icempm
