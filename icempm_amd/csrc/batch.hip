// batch.hip -- many independent indentation cases of uneven size advanced
// together: one launch of the grid update and one of the fused particle pass
// per substep cover every case, and the substep loop is captured once in a
// HIP graph and replayed.
//
// Storage (per batch, cases contiguous, each case in band order):
//   F64 variant    positions: 2 rows of double; vel(2) Fe(4) Bp(4) Jp_inv(1):
//                  11 rows of double
//   F32 variant    positions: 48-bit fixed point over the case's grid width,
//                  3 rows of u32 (x high 32 bits, y high 32 bits, both low
//                  16-bit halves); vel(2) Fe-I(4) Bp(4) Jp_inv-1(1): 11 rows
//                  of float.  Fe and Jp_inv are stored as their deviations
//                  from 1 so that float keeps them to ~1e-12 absolute.
//   crushed        1 byte per point
//   grain          i16 per point, only when a case has GrainVariability != 0
//   grids          two grids of 3 rows of double per node (mass, momentum;
//                  velocity after the update), cases concatenated, swapped
//                  every substep; the F32 variant also keeps one FP32
//                  velocity pair per node for the gathers
// All arithmetic is FP64 in both variants (the return map included); the F32
// variant only stores less: a fused pass moves 49 B per point instead of 97 B.
// A case's NACC strengths are uniform (constants per case) or vary per grain
// id as GetParametersForGrain makes them (gpu_implementation5.cu:97-111).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/icempm.h"
#include "../host/parameters_sim.h"
#include "internal.h"
#include "mlp_device.h"

#pragma clang fp contract(fast)

namespace icempm {
namespace batch {

struct CaseParams {
    double dt, vol, mass, h, h_inv, Dinv, mu, kappa, gravity, indRsq, indV;
    double ics, its, iss, tanPhi, dpThr, x0, y0;
    double pmax, msq, betaPmax, onePlus2Beta, slope;   // NACC constants of the case's material (grain 0)
    double grainVar;        // GrainVariability: 0 = the constants above hold for every point
    // FP32 storage: a coordinate is a 48-bit integer u with x / h = u 2^-F,
    // F = 48 - the bits a cell index of the case needs: the quantum in metres
    // (h 2^-F) and its inverse, and in cell units 2^-F, 2^(16 - F), 2^F
    double posQ, posInvQ, cellQ, cellQ16, cellInvQ;
    double gxInv;           // 1 / gx: a node's row from its index
    double scm;             // dt V_p D^-1 / m_p: the stress term of the affine momentum per unit mass
    int gx, gy, gridTotal, nPts, ptOffset, gridOffset;
    int nbins, binOffset;   // indenter force bins: n_indenter_subdivisions, first double in the bin array
};

#ifndef ICEMPM_BATCH_CHAINS
#define ICEMPM_BATCH_CHAINS 2   // independent chains of a batched run (groups of cases)
#endif
constexpr int kWX = 48, kWS = 52, kWY = 4;   // window width, padded row stride, rows

// Point rows.  FP64 storage: positions FP64.  FP32 storage: positions as
// 48-bit fixed point in units of h 2^-F, the cell index in the top 48 - F
// bits (F = 48 - the bits of the larger grid dimension: a quantum of at most
// ~2 gx h / 2^48, ~2e-14 m for a 3 m grid; a substep moves a point by ~1e-6
// m, which an FP32 coordinate could not resolve, and 32 bits -- a quantum of
// ~7e-10 m -- rounded each displacement by ~1e-4 of itself, enough to move
// the force history of a fracturing case by ~1%); the base cell and in-cell
// offset of a stencil are exact functions of the integer.  The other values
// FP32.
template <bool F32>
struct Rows {
    using V = typename std::conditional<F32, float, double>::type;
    using P = typename std::conditional<F32, unsigned int, double>::type;
    // grid velocity as the gathers read it: FP64 rows 1-2 of the grid, or
    // (FP32 storage) one FP32 pair per node written by the grid update
    using G = typename std::conditional<F32, float2, double>::type;
    // FP64: rows x, y.  FP32: one 16-byte record per slot, {high 32 bits of
    // x, of y, the two low 16-bit halves, the crushed flag | grain id << 16}
    // (one load and one store per pass; the crushed and grain rows are then
    // unused)
    static constexpr int kPosRows = F32 ? 4 : 2;
    P *pos;            // kPosRows values per slot
    V *rest;           // 11 rows: vel(2) Fe(4) Bp(4) Jp(1)
    unsigned char *crushed;
    short *grain;      // null when every case is uniform
    size_t pitch;
};

constexpr double kFix48Max = 281474976710655.0;   // 2^48 - 1

// Element i of a row through a 32-bit byte offset from the row's start: the
// row start is wavefront-uniform (scalar registers), so the access takes the
// scalar-base + 32-bit vector-offset form, with no 64-bit address arithmetic
// per lane.  Rows hold < 2^29 elements (finalize refuses larger batches), so
// i * sizeof(T) < 2^32.
template <class T>
static __device__ __forceinline__ T *at(T *row, unsigned i) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(row) + i * (unsigned)sizeof(T));
}

// FP32 storage: a coordinate as 48-bit fixed point, u = hi 2^16 + lo, in
// metres (round to nearest; below 0 to 0, above 2^48 - 1 quanta and NaN to
// 2^48 - 1, past the grid's edge, where the next stencil check flags them).
// Every step is exact in FP64 except the scalings by posQ / posInvQ.
static __host__ __device__ inline double fix48_get(unsigned int hi, unsigned int lo, const CaseParams &cp) {
    return ((double)hi * 65536.0 + (double)lo) * cp.posQ;
}
static __host__ __device__ inline void fix48_put(double x, const CaseParams &cp, unsigned int &hi, unsigned int &lo) {
    double u = rint(x * cp.posInvQ);
    u = (u >= 0.0 && u <= kFix48Max) ? u : (u < 0.0 ? 0.0 : kFix48Max);
    const double uh = floor(u * (1.0 / 65536.0));
    hi = (unsigned int)uh;
    lo = (unsigned int)(u - uh * 65536.0);
}
// The same in cell units (xr = x / h = u 2^-F): exact both ways, the
// scalings being powers of 2.  fix_put_cells returns u.
static __device__ __forceinline__ double fix_cells(unsigned int hi, unsigned int lo, const CaseParams &cp) {
    return fma((double)hi, cp.cellQ16, (double)lo * cp.cellQ);
}
static __device__ __forceinline__ double fix_put_cells(double xr, const CaseParams &cp, unsigned int &hi,
                                                       unsigned int &lo) {
    double u = rint(xr * cp.cellInvQ);
    u = (u >= 0.0 && u <= kFix48Max) ? u : (u < 0.0 ? 0.0 : kFix48Max);
    const double uh = floor(u * (1.0 / 65536.0));
    hi = (unsigned int)uh;
    lo = (unsigned int)(u - uh * 65536.0);
    return u;
}
// The quadratic B-spline stencil (gpu_implementation5.cu:185-193) of a
// coordinate in cell units: base cell floor(xr - 0.5), 0 below xr = 0.5 (as
// the truncation of the FP64 form gives), and offset fx = xr - base, both
// exact for the xr of a record.
static __device__ __forceinline__ void stencil_cells(double xr, int &base, double &fx, double &w0, double &w1,
                                                     double &w2) {
    const double b = fmax(floor(xr - 0.5), 0.0);
    base = (int)b;
    fx = xr - b;
    const double v0 = 1.5 - fx, v1 = fx - 1.0, v2 = fx - 0.5;
    w0 = 0.5 * v0 * v0;
    w1 = 0.75 - v1 * v1;
    w2 = 0.5 * v2 * v2;
}

// Point position of slot pt (both storage formats) as FP64; store it back,
// and the value the stored form holds.
template <bool NT>
static __device__ __forceinline__ void pos_load(const double *pos, size_t pp, int pt, const CaseParams &,
                                                double &x, double &y) {
    const unsigned u = (unsigned)pt;
    if constexpr (NT) {
        x = __builtin_nontemporal_load(at(pos, u));
        y = __builtin_nontemporal_load(at(pos + pp, u));
    } else {
        x = *at(pos, u);
        y = *at(pos + pp, u);
    }
}
typedef unsigned int prec4 __attribute__((ext_vector_type(4)));
// FP32 storage: the slot's record, its coordinates in metres; returns its
// last word (crushed flag in bits 0-7, grain id in bits 16-31)
template <bool NT>
static __device__ __forceinline__ unsigned pos_load(const unsigned int *pos, size_t, int pt, const CaseParams &cp,
                                                    double &x, double &y) {
    const prec4 *rec = reinterpret_cast<const prec4 *>(pos) + pt;
    prec4 v;
    if constexpr (NT)
        v = __builtin_nontemporal_load(rec);
    else
        v = *rec;
    x = fix48_get(v.x, v.z & 0xFFFFu, cp);
    y = fix48_get(v.y, v.z >> 16, cp);
    return v.w;
}
// FP64 storage: stores (x, y), through the L2 (a wavefront's 512 B of a row
// straddle cache lines that the neighbouring wavefronts complete: measured
// 6% faster than non-temporal stores).  The FP32 record streams
// non-temporally (measured faster that way).
static __device__ __forceinline__ void pos_store(double *pos, size_t pp, int pt, double x, double y) {
    const unsigned u = (unsigned)pt;
    *at(pos, u) = x;
    *at(pos + pp, u) = y;
}

// FP32 storage keeps a point's Fe - I as one 16-byte record in the space of
// rest rows 2-5 (one float4 per point, one load and one store per pass
// instead of four of each); FP64 storage keeps the four rows.
typedef float fe4 __attribute__((ext_vector_type(4)));
static __device__ __forceinline__ fe4 *fe_rec(float *R, size_t pp) { return reinterpret_cast<fe4 *>(R + pp * 2); }
static __device__ __forceinline__ const fe4 *fe_rec(const float *R, size_t pp) {
    return reinterpret_cast<const fe4 *>(R + pp * 2);
}

// the velocity of a node (index within its case's grid) as the gathers read it
static __device__ __forceinline__ void grid_v(const double *g, int node, size_t gpitch, double &v0, double &v1) {
    v0 = g[gpitch + node];
    v1 = g[2 * gpitch + node];
}
static __device__ __forceinline__ void grid_v(const float2 *g, int node, size_t, float &v0, float &v1) {
    const float2 a = *at(g, (unsigned)node);
    v0 = a.x;
    v1 = a.y;
}

static __device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Band order of a case's base cells: the rows in pairs (bands, a last odd
// row alone), a band column by column with the two cells of a column
// together.  Consecutive cells then cover two rows: a wavefront's 64 points
// (16 cells at 4 per cell) lie in 2 x 8 cells, whose stencils touch 4 x 10
// grid nodes, where 1 x 16 cells of a row-major order touch 3 x 18.  The
// keys of a case are a bijection onto [0, gx * gy).
static __host__ __device__ inline int band_key(int x, int y, int gx, int gy) {
    const int band = y >> 1;
    const int rows = (gy - 2 * band) < 2 ? 1 : 2;
    return band * 2 * gx + x * rows + (y & 1);
}
static __host__ __device__ inline void band_cell(int key, int gx, int gy, int &x, int &y) {
    const int band = key / (2 * gx);
    const int w = key - band * 2 * gx;
    const int rows = (gy - 2 * band) < 2 ? 1 : 2;
    x = w / rows;
    y = 2 * band + w % rows;
}

static __device__ __forceinline__ KParams kp_of(const CaseParams &c, int *err) {
    KParams k{};
    k.err = err;
    k.gridX = c.gx;
    k.gridY = c.gy;
    k.dt = c.dt; k.vol = c.vol; k.mass = c.mass; k.h = c.h; k.h_inv = c.h_inv; k.Dinv = c.Dinv;
    k.mu = c.mu; k.kappa = c.kappa; k.gravity = c.gravity; k.indRsq = c.indRsq; k.indV = c.indV;
    k.ics = c.ics; k.its = c.its; k.iss = c.iss; k.tanPhi = c.tanPhi; k.dpThr = c.dpThr;
    k.grainVar = c.grainVar;
    return k;
}

// Return map of crushed points: the exact Drucker-Prager map, the exact map
// while recording (trial, projected) pairs, or the learned map with the
// exact one as fallback.
enum { kMapExact = 0, kMapHarvest = 1, kMapLearned = 2 };

struct LearnArgs {
    const unsigned short *wbf;   // MLP weights, BF16 bits
    const float *pf;             // MLP parameters, FP32 (biases)
    mlpdev::Norm norm;
    double tol;                  // accepted when the yield residual is <= tol
    float *hx, *ht;              // harvest: inputs (n x 5), targets (n x 3)
    unsigned long long *hcount;
    int hcap;
    // [accepted, fell back, inside the surface, tension accepted, tension fell
    // back, shear accepted, shear fell back] (the last four: map kernel only)
    unsigned long long *stats;
    int branches;                // kMapTension | kMapShear: branches the network proposes for
};

constexpr int kMapStats = 7;
enum { kMapTension = 1, kMapShear = 2 };

// A crushed point whose projection is deferred to the map kernel: its trial
// elastic deformation (FP64) and Jp_inv, and where it lives.
struct MapEntry {
    int pt, cs;
    double F[4];
    double Jp;
    double x[2], v[2], B[4];   // position, velocity and Bp for its P2G
};

// Queue of deferred points of one pass (learned map only).
struct MapQueue {
    MapEntry *q;
    unsigned int *count;   // appended by the particle pass
    int cap;
};

// The branch of the exact map (Wolper_Drucker_Prager, gpu_implementation5.cu:48-94)
// a crushed trial state falls in, decided exactly: the tension branch, the
// shear projection onto q = q_lim, or inside the surface (the map is the
// identity there).  qlim is set for the shear branch.
enum { kBrInside = 0, kBrTension = 1, kBrShear = 2 };
static __device__ __forceinline__ int map_branch(const KParams &k, const Trial &r, double Jp, double &qlim) {
    qlim = 0;
    if (r.p < -k.dpThr || Jp < 1) return kBrTension;
    qlim = (r.p > k.ics) ? 0.0 : fmin((r.p + k.dpThr) * k.tanPhi, k.iss);
    return (r.q >= qlim) ? kBrShear : kBrInside;
}

// Raw features of a crushed point for the map: log s0, log s1, log Jp_inv,
// a hardening variable the reference does not have (0), and the yield-state
// flag: 1 in the tension branch, 0 in the shear branch.
static __device__ __forceinline__ void map_features(const Trial &r, double Jp, int br, double f[5]) {
    f[0] = log(r.svd.s0);
    f[1] = log(r.svd.s1);
    f[2] = log(Jp);
    f[3] = 0.0;
    f[4] = (br == kBrTension) ? 1.0 : 0.0;
}

// Yield residual of a proposed projected state (s0, s1, Jpn), as stress
// errors relative to the case's strengths:
//  tension  the state must be isotropic at p = -p_thr:
//           max(|p(s) + p_thr| / pmax, q(s) / qmax, |Jpn / Jp_exact - 1|)
//  shear    the deviator (scaled with the trial J, as the map scales it)
//           must sit at q_lim and the mean of s^2 must be kept:
//           max(|q~(s) - q_lim| / qmax, kappa |mean s^2 - mean s_tr^2| / pmax)
static __device__ __forceinline__ double yield_residual(const KParams &k, const Trial &r, double Jp, int br,
                                                        double qlim, double s0, double s1, double Jpn) {
    const double a = s0 * s0, b = s1 * s1;
    if (br == kBrTension) {
        const double Je_new = sqrt(-2. * -k.dpThr / k.kappa + 1.);
        const double Jn = s0 * s1;
        const double pn = -(k.kappa / 2.) * (Jn * Jn - 1.);
        const double qn = (k.mu / Jn) * fabs(a - b);
        double res = fabs(pn + k.dpThr) / k.ics;
        res = fmax(res, qn / k.iss);
        res = fmax(res, fabs(Jpn / (Jp * (Je_new / r.Je)) - 1.));
        return res;
    }
    const double qt = (k.mu / r.Je) * fabs(a - b);
    const double dm = 0.5 * (a + b) - 0.5 * (r.ss0 + r.ss1);
    return fmax(fabs(qt - qlim) / k.iss, k.kappa * fabs(dm) / k.ics);
}

// Bytes of the per-wavefront P2G window.
constexpr int kWaveWinBytes = 3 * kWY * kWS * 8;

// Fused particle pass over all cases (G2P of substep k and/or P2G of k+1).
// The chunk table, the case table and the point rows lead the arguments,
// which the kernel receives preloaded in SGPRs: a workgroup's first load is
// its chunk {first point, end, case}, the point loads follow it at once,
// and the case's parameters arrive meanwhile.
// Quad-level P2G.  A DPP read of a double from another lane of the quad
// (quad_perm control CTRL; every lane of the quad must be active, so no
// lane keeps an old value and none is needed)
template <int CTRL>
static __device__ __forceinline__ double quad_dpp(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kQuadXor3 = 0x1B;   // quad_perm 1032, 2301, 3210
// whether the four lanes of this lane's quad hold the same key (all lanes
// of the wavefront active)
static __device__ __forceinline__ bool quad_uniform(unsigned key) {
    const int k = (int)key;
    return k == __builtin_amdgcn_mov_dpp(k, kQuadXor1, 0xF, 0xF, true) &&
           k == __builtin_amdgcn_mov_dpp(k, kQuadXor2, 0xF, 0xF, true) &&
           k == __builtin_amdgcn_mov_dpp(k, kQuadXor3, 0xF, 0xF, true);
}
// One component's ten node values of a quad lane (step 1 layout, below)
// summed over the quad and added to the window row Wc.
static __device__ __forceinline__ void quad_component(int lane, const double (&K)[5], const double (&D)[4], int eA,
                                                      int eB, int eC, double *Wc) {
    const int sy = (lane >> 1) & 1;
    double k[5];
#pragma unroll
    for (int s = 0; s < 4; ++s) k[s] = K[s] + quad_dpp<kQuadXor1>(D[s]);
    k[4] = K[4] + quad_dpp<kQuadXor1>(K[4]);
    // step 2: the partners across bit 1 split slots (0, 1) and (2, 3)
    const double x01 = sy ? k[0] : k[1], y01 = sy ? k[1] : k[0];
    const double x23 = sy ? k[2] : k[3], y23 = sy ? k[3] : k[2];
    atomicAdd(&Wc[eA], y01 + quad_dpp<kQuadXor2>(x01));
    atomicAdd(&Wc[eB], y23 + quad_dpp<kQuadXor2>(x23));
    const double k4 = k[4] + quad_dpp<kQuadXor2>(k[4]);
    if ((lane & 3) == 0) atomicAdd(&Wc[eC], k4);
}
// The P2G of a quad whose four points share their base cell, into the
// window (e0: the base's element): node (i, j) of the stencil gets
// W (v + A d), W = wx_i wy_j, d = ((i, j) - f) h, summed over the four
// points.  Step 1 pairs lanes across lane bit 0 (sx): slot s < 3 holds node
// (2 sx, s), kept, and (2 - 2 sx, s), sent to the partner; slot 3 node
// (1, sx), kept, and (1, 1 - sx), sent; slot 4 node (1, 2) on both.  After
// it a lane holds pair sums of five (sx = 0) or four (sx = 1) nodes.  Step 2
// pairs lanes across bit 1 (sy): of slots 0 and 1 the lane keeps slot sy,
// of slots 2 and 3 slot 2 + sy, and lane 0 of the quad keeps slot 4.  Each
// lane then adds two or three of the nine node sums per row: 6-9 LDS
// atomics instead of 27, the nodes of one instruction still distinct
// within the quad.
static __device__ __forceinline__ void quad_scatter(int lane, int e0, double h, double fx0, double fx1, double wx0,
                                                    double wx1, double wx2, double wy0, double wy1, double wy2,
                                                    double vx, double vy, double A00, double A01, double A10,
                                                    double A11, double *W0, double *W1, double *W2) {
    const int sx = lane & 1, sy = (lane >> 1) & 1;
    const double dx0 = (0.0 - fx0) * h, dx1 = (1.0 - fx0) * h, dx2 = (2.0 - fx0) * h;
    const double dy0 = (0.0 - fx1) * h, dy1 = (1.0 - fx1) * h, dy2 = (2.0 - fx1) * h;
    const double xa0 = vx + A00 * dx0, xa1 = vx + A00 * dx1, xa2 = vx + A00 * dx2;
    const double xb0 = vy + A10 * dx0, xb1 = vy + A10 * dx1, xb2 = vy + A10 * dx2;
    const double ya0 = A01 * dy0, ya1 = A01 * dy1, ya2 = A01 * dy2;
    const double yb0 = A11 * dy0, yb1 = A11 * dy1, yb2 = A11 * dy2;
    // the kept and sent columns of slots 0-2, and the kept and sent rows of slot 3
    const double wxK = sx ? wx2 : wx0, xaK = sx ? xa2 : xa0, xbK = sx ? xb2 : xb0;
    const double wxD = sx ? wx0 : wx2, xaD = sx ? xa0 : xa2, xbD = sx ? xb0 : xb2;
    const double wyK = sx ? wy1 : wy0, yaK = sx ? ya1 : ya0, ybK = sx ? yb1 : yb0;
    const double wyD = sx ? wy0 : wy1, yaD = sx ? ya0 : ya1, ybD = sx ? yb0 : yb1;
    const double wy[3] = {wy0, wy1, wy2}, ya[3] = {ya0, ya1, ya2}, yb[3] = {yb0, yb1, yb2};
    double WK[5], WD[4];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        WK[s] = wxK * wy[s];
        WD[s] = wxD * wy[s];
    }
    WK[3] = wx1 * wyK;
    WD[3] = wx1 * wyD;
    WK[4] = wx1 * wy2;
    // the nodes a lane adds: slot sy's node (2 sx, sy); slot 2 + sy's node,
    // (2 sx, 2) or (1, sx); slot 4's node (1, 2)
    const int eA = e0 + 2 * sx + sy * kWS;
    const int eB = e0 + (sy ? 1 + sx * kWS : 2 * sx + 2 * kWS);
    const int eC = e0 + 1 + 2 * kWS;
    quad_component(lane, WK, WD, eA, eB, eC, W0);
    double K[5], D[4];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        K[s] = WK[s] * (xaK + ya[s]);
        D[s] = WD[s] * (xaD + ya[s]);
    }
    K[3] = WK[3] * (xa1 + yaK);
    D[3] = WD[3] * (xa1 + yaD);
    K[4] = WK[4] * (xa1 + ya2);
    quad_component(lane, K, D, eA, eB, eC, W1);
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        K[s] = WK[s] * (xbK + yb[s]);
        D[s] = WD[s] * (xbD + yb[s]);
    }
    K[3] = WK[3] * (xb1 + ybK);
    D[3] = WD[3] * (xb1 + ybD);
    K[4] = WK[4] * (xb1 + yb2);
    quad_component(lane, K, D, eA, eB, eC, W2);
}

template <bool F32, bool G2P, bool P2G, int RMAP>
__global__ __launch_bounds__(256) void k_batch_particles(const int4 *__restrict__ chunks, int nChunks,
                                                         const CaseParams *__restrict__ cases,
                                                         typename Rows<F32>::P *pos,
                                                         typename Rows<F32>::V *rest, unsigned char *crushedRow,
                                                         size_t pitch, const typename Rows<F32>::G *__restrict__ gridV,
                                                         double *__restrict__ gridOut, size_t gpitch,
                                                         int *__restrict__ err, const short *__restrict__ grains,
                                                         LearnArgs la, MapQueue mq, int storeVB) {
    const Rows<F32> S{pos, rest, crushedRow, nullptr, pitch};
    __shared__ __attribute__((aligned(16))) double smem[4][kWaveWinBytes / 8];
    const int per = (nChunks + 7) >> 3;
    const int c = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (c >= nChunks) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int4 ch = chunks[c];
    const int beg = ch.x, end = ch.y;
    asm volatile("" ::"s"(ch.z));   // all of the chunk in one load, before the branch
    // FP32 storage: a wavefront issues its point loads and gathers, and later
    // its window flush, at raised priority, computes at the base one
    // (measured: -2.6% per substep; FP64 storage measured slower with it)
    if constexpr (F32) __builtin_amdgcn_s_setprio(2);
    if (beg + wv * 64 >= end) return;
    const CaseParams cp = cases[ch.z];
    const KParams k = kp_of(cp, err);
    const int pt = beg + (int)threadIdx.x;
    const bool valid = pt < end;
    const size_t pp = S.pitch;
    const int gx = cp.gx, gy = cp.gy;
    const auto *__restrict__ gV = gridV + cp.gridOffset;
    double *__restrict__ gO = gridOut + cp.gridOffset;
    using V = typename Rows<F32>::V;
    V *__restrict__ R = S.rest;
    constexpr double fe0 = F32 ? 1.0 : 0.0;   // stored Fe = Fe - fe0 * I, Jp_inv - fe0

    double px = 0, py = 0, vx = 0, vy = 0;
    double F00 = 1, F01 = 0, F10 = 0, F11 = 1;
    double B00 = 0, B01 = 0, B10 = 0, B11 = 0;
    bool defer = false;   // learned map: projection and P2G done by the map kernel
    InvSq rt{};           // the trial invariants, and whether the map left F alone
    bool fresh = false;

    if constexpr (G2P) {
        double Jp = 1;
        unsigned char crushed = 0;
        short grain = 0;
        if (valid) {
            // streamed once per substep: non-temporal, so the grid keeps the L2
            const unsigned u = (unsigned)pt;
            if constexpr (F32) {
                // the record; its coordinates in cell units
                const prec4 r = __builtin_nontemporal_load(reinterpret_cast<const prec4 *>(S.pos) + pt);
                px = fix_cells(r.x, r.z & 0xFFFFu, cp);
                py = fix_cells(r.y, r.z >> 16, cp);
                crushed = (unsigned char)(r.w & 0xFFu);
                grain = (short)(r.w >> 16);
                const fe4 f = __builtin_nontemporal_load(fe_rec(R, pp) + pt);
                F00 = fe0 + (double)f.x;
                F01 = (double)f.y;
                F10 = (double)f.z;
                F11 = fe0 + (double)f.w;
            } else {
                pos_load<true>(S.pos, pp, pt, cp, px, py);
                F00 = fe0 + (double)__builtin_nontemporal_load(at(R + pp * 2, u));
                F01 = (double)__builtin_nontemporal_load(at(R + pp * 3, u));
                F10 = (double)__builtin_nontemporal_load(at(R + pp * 4, u));
                F11 = fe0 + (double)__builtin_nontemporal_load(at(R + pp * 5, u));
                crushed = __builtin_nontemporal_load(at(S.crushed, u));
                if (cp.grainVar != 0) grain = __builtin_nontemporal_load(at(grains, u));
            }
        }
        int bx = 0, by = 0;
        double fx0 = 0, fx1 = 0, wx0, wx1, wx2, wy0, wy1, wy2;
        bool ok;
        if constexpr (F32) {   // cell units: a record always has a stencil
            ok = valid;
            if (ok) {
                stencil_cells(px, bx, fx0, wx0, wx1, wx2);
                stencil_cells(py, by, fx1, wy0, wy1, wy2);
            }
        } else {
            ok = valid && stencil(px, cp.h_inv, bx, fx0, wx0, wx1, wx2) &&
                 stencil(py, cp.h_inv, by, fx1, wy0, wy1, wy2);
            if (valid && !ok) flag_error(k);
        }
        InvSq iv{};
        unsigned char crushed0 = crushed;
        double Jp0 = Jp;
        if (ok) {
            // the gathered velocities stay in their stored type (FP32 pairs
            // or FP64) and widen at each use: fewer live registers
            using GT = typename std::conditional<F32, float, double>::type;
            GT n0[9], n1[9];
            if (bx + 2 < gx && by + 2 < gy) {
                if constexpr (F32) {
                    // a stencil row's three FP32 pairs as one 16-byte load
                    // (8-byte aligned: the unaligned 16-byte access global
                    // memory allows) and one 8-byte load: 6 loads instead of
                    // 9 (measured 2% off the pass)
                    typedef float v4u __attribute__((ext_vector_type(4), aligned(8)));
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const unsigned node = (unsigned)(bx + (by + j) * gx);
                        const v4u a = *reinterpret_cast<const v4u *>(at(gV, node));
                        const float2 c = *at(gV, node + 2);
                        n0[0 * 3 + j] = a.x;
                        n1[0 * 3 + j] = a.y;
                        n0[1 * 3 + j] = a.z;
                        n1[1 * 3 + j] = a.w;
                        n0[2 * 3 + j] = c.x;
                        n1[2 * 3 + j] = c.y;
                    }
                } else {
                    const double *gvx = gV + gpitch + bx + by * gx;
                    const double *gvy = gvx + gpitch;
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) {
                            n0[i * 3 + j] = gvx[i + j * gx];
                            n1[i * 3 + j] = gvy[i + j * gx];
                        }
                }
            } else {
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const int X = bx + i, Y = by + j;
                        n0[i * 3 + j] = 0;
                        n1[i * 3 + j] = 0;
                        if (X >= gx || Y >= gy) {
                            flag_error(k);
                        } else {
                            grid_v(gV, X + Y * gx, gpitch, n0[i * 3 + j], n1[i * 3 + j]);
                        }
                    }
            }
            if constexpr (F32) __builtin_amdgcn_s_setprio(0);
            double v0 = 0, v1 = 0, b00 = 0, b01 = 0, b10 = 0, b11 = 0;
            const double c4 = 4. * cp.h_inv;
            const double wyd0 = wy0 * (0 - fx1), wyd1 = wy1 * (1 - fx1), wyd2 = wy2 * (2 - fx1);
            const double wxs[3] = {wx0, wx1, wx2};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double a0 = n0[i * 3], a1 = n0[i * 3 + 1], a2 = n0[i * 3 + 2];
                const double c0 = n1[i * 3], c1 = n1[i * 3 + 1], c2 = n1[i * 3 + 2];
                const double s0 = wy0 * a0 + wy1 * a1 + wy2 * a2;
                const double s1 = wy0 * c0 + wy1 * c1 + wy2 * c2;
                const double t0 = wyd0 * a0 + wyd1 * a1 + wyd2 * a2;
                const double t1 = wyd0 * c0 + wyd1 * c1 + wyd2 * c2;
                const double wxd = wxs[i] * (i - fx0);
                v0 += wxs[i] * s0;
                v1 += wxs[i] * s1;
                b00 += wxd * s0;
                b10 += wxd * s1;
                b01 += wxs[i] * t0;
                b11 += wxs[i] * t1;
            }
            b00 *= c4; b01 *= c4; b10 *= c4; b11 *= c4;
            const double dt = cp.dt;
            const double dtp = F32 ? dt * cp.h_inv : dt;   // FP32 storage: cell units
            px += v0 * dtp;
            py += v1 * dtp;
            const double M00 = 1.0 + dt * b00, M01 = dt * b01, M10 = dt * b10, M11 = 1.0 + dt * b11;
            const double n00 = M00 * F00 + M01 * F10, n01 = M00 * F01 + M01 * F11;
            const double n10 = M10 * F00 + M11 * F10, n11 = M10 * F01 + M11 * F11;
            F00 = n00; F01 = n01; F10 = n10; F11 = n11;
            // p and q^2 of the trial state (q itself is never needed here:
            // every test compares it with a non-negative limit)
            iv = invariants_sq(F00, F01, F10, F11, cp.mu, cp.kappa);
            crushed0 = crushed;
            Jp0 = Jp;
            if (crushed == 0) {
                bool f;
                if (cp.grainVar != 0) {   // GetParametersForGrain's strengths (case-uniform branch)
                    f = fails_sq(k, grain, iv.p, iv.q2);
                } else if (iv.p < 0) {
                    // tension line q > slope (p + 3e6), slope >= 0 and p + 3e6 >= 0 here
                    const double ql = cp.slope * (iv.p - -3e6);
                    f = (iv.p < -3e6) || (iv.q2 > ql * ql);
                } else {
                    f = cp.onePlus2Beta * iv.q2 + cp.msq * (iv.p + cp.betaPmax) * (iv.p - cp.pmax) > 0;
                }
                crushed = f ? 1 : 0;
            }
            vx = v0; vy = v1;
            B00 = b00; B01 = b01; B10 = b10; B11 = b11;
        }
        // ---- return map (convergent: the learned map runs wavefront-wide) ----
        // The trial SVD only for crushed points (few: the wavefront mostly
        // skips it); p and q for the NACC test came from the invariants.
        // Jp_inv matters only to the map, so only crushed points load it.
        const bool crushedOk = ok && crushed == 1;
        if (crushedOk) {
            Jp = fe0 + (double)*at(R + pp * 10, (unsigned)pt);
            Jp0 = Jp;
        }
        if constexpr (RMAP == kMapExact) {
            if (crushedOk && dp_projects_sq(k, iv.p, iv.q2, Jp))
                dp_return(k, compute_pq_fast(F00, F01, F10, F11, cp.mu, cp.kappa), F00, F01, F10, F11, Jp);
        } else if constexpr (RMAP == kMapHarvest) {
            Trial r{};
            if (crushedOk) r = compute_pq_fast(F00, F01, F10, F11, cp.mu, cp.kappa);
            double f[5] = {0, 0, 0, 0, 0}, t[3] = {0, 0, 0};
            bool usable = false;
            if (crushedOk) {
                double qlim;
                const int br = map_branch(k, r, Jp, qlim);
                map_features(r, Jp, br, f);
                dp_return(k, r, F00, F01, F10, F11, Jp);
                if (br != kBrInside) {
                    const Trial rn = compute_pq(F00, F01, F10, F11, cp.mu, cp.kappa);
                    usable = r.svd.s1 > 0 && rn.svd.s1 > 0;
                    t[0] = log(rn.svd.s0) - f[0];
                    t[1] = log(rn.svd.s1) - f[1];
                    t[2] = log(Jp) - f[2];
                }
            }
            // one counter update per wavefront
            const unsigned long long um = __ballot(usable);
            if (um) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(la.hcount, (unsigned long long)__popcll(um));
                base = __shfl(base, 0, 64);
                const unsigned long long slot = base + (unsigned long long)__popcll(um & ((1ull << lane) - 1ull));
                if (usable && slot < (unsigned long long)la.hcap) {
                    for (int q = 0; q < 5; ++q) la.hx[slot * 5 + q] = (float)f[q];
                    for (int q = 0; q < 3; ++q) la.ht[slot * 3 + q] = (float)t[q];
                }
            }
        } else {
            // the branch is decided exactly; inside the surface the map is
            // the identity; points that project are queued for the map
            // kernel (network proposal, residual check, exact fallback) unless
            // their state is inverted, which takes the exact map here
            double qlim = 0;
            Trial r{};
            if (crushedOk) r = compute_pq_fast(F00, F01, F10, F11, cp.mu, cp.kappa);
            const int br = crushedOk ? map_branch(k, r, Jp, qlim) : kBrInside;
            const bool proj = crushedOk && br != kBrInside;
            // the network proposes only in the branches it is enabled for
            // (the autotuner switches off a branch whose proposals the yield
            // gate does not pass); the others take the exact map here
            const int bit = (br == kBrTension) ? kMapTension : kMapShear;
            defer = proj && r.svd.s1 > 0 && (la.branches & bit) != 0;
            if (proj && !defer) dp_return(k, r, F00, F01, F10, F11, Jp);
            const unsigned long long dm = __ballot(defer);
            if (dm) {
                unsigned int base = 0;
                if (lane == 0) base = atomicAdd(mq.count, (unsigned int)__popcll(dm));
                base = __shfl(base, 0, 64);
                const unsigned int slot = base + (unsigned int)__popcll(dm & ((1ull << lane) - 1ull));
                if (defer) {
                    if (slot < (unsigned int)mq.cap) {
                        MapEntry &e = mq.q[slot];
                        e.pt = pt;
                        e.cs = ch.z;
                        e.F[0] = F00; e.F[1] = F01; e.F[2] = F10; e.F[3] = F11;
                        e.Jp = Jp;
                        // the position at its stored precision, as the P2G here would take it
                        if constexpr (F32) {
                            unsigned int hi, lo;
                            e.x[0] = fix_put_cells(px, cp, hi, lo) * cp.posQ;
                            e.x[1] = fix_put_cells(py, cp, hi, lo) * cp.posQ;
                        } else {
                            e.x[0] = px;
                            e.x[1] = py;
                        }
                        e.v[0] = vx; e.v[1] = vy;
                        e.B[0] = B00; e.B[1] = B01; e.B[2] = B10; e.B[3] = B11;
                    } else {
                        flag_error(k);
                    }
                }
            }
            const unsigned long long fb = __ballot(proj && !defer), ins = __ballot(crushedOk && !proj);
            if (lane == 0 && (fb | ins)) {
                if (fb) atomicAdd(&la.stats[1], (unsigned long long)__popcll(fb));
                if (ins) atomicAdd(&la.stats[2], (unsigned long long)__popcll(ins));
            }
        }
        // a point the map left alone keeps the stress of its trial state: the
        // P2G takes it from the invariants (with FP32 storage they describe
        // F before its rounding to the stored FP32, which differs from it by
        // ~6e-8 of Fe - I)
        rt = iv;
        fresh = ok && crushed == 0;
        if (ok) {
            // px, py become what the rows hold: the P2G takes them at their
            // stored precision
            if constexpr (F32) {
                // a position outside the grid (or NaN) is an error here, as
                // the next stencil of FP64 rows would report it; the fixed
                // point clamps it
                if (!(px >= 0.0 && px < (double)gx && py >= 0.0 && py < (double)gy)) flag_error(k);
                unsigned int hx, lx, hy, ly;
                px = fix_put_cells(px, cp, hx, lx) * cp.cellQ;
                py = fix_put_cells(py, cp, hy, ly) * cp.cellQ;
                const prec4 r = {hx, hy, lx | (ly << 16),
                                 (unsigned)crushed | ((unsigned)(unsigned short)grain << 16)};
                __builtin_nontemporal_store(r, reinterpret_cast<prec4 *>(S.pos) + pt);
            } else {
                pos_store(S.pos, pp, pt, px, py);
            }
            const unsigned u = (unsigned)pt;
            if (!defer) {
                if constexpr (F32) {
                    const fe4 f = {(float)(F00 - fe0), (float)F01, (float)F10, (float)(F11 - fe0)};
                    __builtin_nontemporal_store(f, fe_rec(R, pp) + pt);
                } else {   // as pos_store
                    *at(R + pp * 2, u) = (V)F00;
                    *at(R + pp * 3, u) = (V)F01;
                    *at(R + pp * 4, u) = (V)F10;
                    *at(R + pp * 5, u) = (V)F11;
                }
            }
            // v and Bp: a fused pass uses them at once (a deferred point's go
            // to the map kernel in its queue entry) and the next G2P
            // recomputes them before any use; the last pass of a run stores
            // them (storeVB), for readback, checkpoints and a P2G that
            // starts afresh
            if (!P2G || storeVB) {
                R[pt] = (V)vx;
                R[pp + pt] = (V)vy;
                R[pp * 6 + pt] = (V)B00;
                R[pp * 7 + pt] = (V)B01;
                R[pp * 8 + pt] = (V)B10;
                R[pp * 9 + pt] = (V)B11;
            }
            if (Jp != Jp0 && !defer) *at(R + pp * 10, u) = (V)(Jp - fe0);
            if (!F32 && crushed != crushed0) *at(S.crushed, u) = crushed;   // FP32: in the record
            if constexpr (F32) {
                // Fe is what the next substep reads back: the P2G takes it at
                // its stored precision.  v and Bp stay FP64: the next G2P
                // recomputes them (what the last pass of a run stores is
                // read only by a fresh P2G and the readback).
                F00 = fe0 + (double)(V)(F00 - fe0); F01 = (double)(V)F01;
                F10 = (double)(V)F10; F11 = fe0 + (double)(V)(F11 - fe0);
            }
        }
    }

    if constexpr (P2G) {
        bool have = valid;
        if constexpr (!G2P) {
            if (valid) {
                if constexpr (F32) {
                    const prec4 r = reinterpret_cast<const prec4 *>(S.pos)[pt];
                    px = fix_cells(r.x, r.z & 0xFFFFu, cp);
                    py = fix_cells(r.y, r.z >> 16, cp);
                } else {
                    pos_load<false>(S.pos, pp, pt, cp, px, py);
                }
                vx = (double)R[pt];
                vy = (double)R[pp + pt];
                if constexpr (F32) {
                    const fe4 f = fe_rec(R, pp)[pt];
                    F00 = fe0 + (double)f.x;
                    F01 = (double)f.y;
                    F10 = (double)f.z;
                    F11 = fe0 + (double)f.w;
                } else {
                    F00 = fe0 + (double)R[pp * 2 + pt];
                    F01 = (double)R[pp * 3 + pt];
                    F10 = (double)R[pp * 4 + pt];
                    F11 = fe0 + (double)R[pp * 5 + pt];
                }
                B00 = (double)R[pp * 6 + pt];
                B01 = (double)R[pp * 7 + pt];
                B10 = (double)R[pp * 8 + pt];
                B11 = (double)R[pp * 9 + pt];
            }
        } else {
            have = valid && !defer && (F32 || !(px != px || py != py));
        }
        int bx = 0, by = 0;
        double fx0 = 0, fx1 = 0, wx0, wx1, wx2, wy0, wy1, wy2;
        bool ok;
        if constexpr (F32) {   // cell units
            ok = have;
            if (ok) {
                stencil_cells(px, bx, fx0, wx0, wx1, wx2);
                stencil_cells(py, by, fx1, wy0, wy1, wy2);
            }
        } else {
            ok = have && stencil(px, cp.h_inv, bx, fx0, wx0, wx1, wx2) &&
                 stencil(py, cp.h_inv, by, fx1, wy0, wy1, wy2);
        }
        if (valid && !ok && !defer) flag_error(k);   // deferred points scatter in the map kernel
        // the box of base cells of the lanes whose stencils can reach the
        // window (a lane outside the grid touches no node): the window's
        // origin and the w x hgt nodes the stencils touch, which alone are
        // cleared and flushed; lanes map onto the box 16, 32 or 64 columns
        // at a time
        const bool inb = ok && bx < gx && by < gy;
        const bool any = __ballot(inb) != 0;
        const unsigned lx = wave_min_u32(inb ? (unsigned)bx : 0xFFFFFFFFu);
        const unsigned ly = wave_min_u32(inb ? (unsigned)by : 0xFFFFFFFFu);
        const unsigned hx = wave_max_u32(inb ? (unsigned)bx : 0u);
        const unsigned hy = wave_max_u32(inb ? (unsigned)by : 0u);
        const int ox = any ? (int)lx : INT_MAX;
        const int oy = any ? (int)ly : INT_MAX;
        const int bw = any ? min((int)hx + 3 - ox, kWX) : 0;
        const int bh = any ? min((int)hy + 3 - oy, kWY) : 0;
        const int bsh = bw <= 16 ? 4 : (bw <= 32 ? 5 : 6);
        const int bcol = lane & ((1 << bsh) - 1), brow0 = lane >> bsh, bstep = 64 >> bsh;
        double *W0 = smem[wv];
        double *W1 = W0 + kWY * kWS, *W2 = W1 + kWY * kWS;
        if (bcol < bw) {
            for (int ly = brow0; ly < bh; ly += bstep) {
                const int e = ly * kWS + bcol;
                W0[e] = 0;
                W1[e] = 0;
                W2[e] = 0;
            }
        }
        wsync();
        // A quad of lanes whose four points share one base cell (in band
        // order most quads, until points cross into other cells) and whose
        // stencil lies in the window sums its contributions in registers
        // before they reach the window.  Every lane is active here.
        const bool winFits = ok && bx + 2 < gx && by + 2 < gy && (unsigned)(bx - ox) < (unsigned)(kWX - 2) &&
                             (unsigned)(by - oy) < (unsigned)(kWY - 2);
        const bool quad = quad_uniform(winFits ? (unsigned)bx | (unsigned)by << 16 : 0xFFFF0000u | (unsigned)lane);
        if (ok) {
            double t00, t01, t11;
            if (fresh)
                kirchhoff_from_inv(rt, t00, t01, t11);
            else
                kirchhoff_rcp(F00, F01, F10, F11, cp.mu, cp.kappa, t00, t01, t11);
            // Every point of a case has the same mass m: the window sums the
            // contributions per unit mass, W (v + (Bp - m^-1 dt V D^-1 tau) d),
            // and the flush multiplies by m once per node.
            const double m = cp.mass, h = cp.h;
            const double A00 = B00 - cp.scm * t00, A01 = B01 - cp.scm * t01;
            const double A10 = B10 - cp.scm * t01, A11 = B11 - cp.scm * t11;
            const int lx0 = bx - ox, ly0 = by - oy;
            // 0 <= lx0 <= kWX - 3 and 0 <= ly0 <= kWY - 3, one unsigned
            // compare each (the origin is the least base cell of the lanes
            // that get here, so lx0, ly0 >= 0 already)
            if (quad) {
                quad_scatter(lane, ly0 * kWS + lx0, h, fx0, fx1, wx0, wx1, wx2, wy0, wy1, wy2, vx, vy, A00, A01, A10,
                             A11, W0, W1, W2);
            } else if (winFits) {
                const int e0 = ly0 * kWS + lx0;
                // The four lanes of a quad (in band order mostly the points of
                // one cell) walk the 3 x 3 stencil from four different nodes,
                // so their LDS atomics never meet at one address: lane bit 0
                // shifts the columns, bit 1 the rows, by one.  Step k of an
                // axis visits node (k + s) mod 3 = s, 1 + s, 2 - 2 s; only the
                // weights need selecting, the offsets follow from the index.
                const int sx = lane & 1, sy = (lane >> 1) & 1;
                const int xe[3] = {sx, 1 + sx, 2 - 2 * sx};
                const int iy[3] = {sy, 1 + sy, 2 - 2 * sy};
                const int ye[3] = {iy[0] * kWS, iy[1] * kWS, iy[2] * kWS};
                const double xw[3] = {sx ? wx1 : wx0, sx ? wx2 : wx1, sx ? wx0 : wx2};
                const double yw[3] = {sy ? wy1 : wy0, sy ? wy2 : wy1, sy ? wy0 : wy2};
                double xa[3], xb[3], ya[3], yb[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double dx = ((double)xe[q] - fx0) * h, dy = ((double)iy[q] - fx1) * h;
                    xa[q] = vx + A00 * dx;
                    xb[q] = vy + A10 * dx;
                    ya[q] = A01 * dy;
                    yb[q] = A11 * dy;
                }
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const double W = xw[a] * yw[b];
                        const int e = e0 + xe[a] + ye[b];
                        atomicAdd(&W0[e], W);
                        atomicAdd(&W1[e], W * (xa[a] + ya[b]));
                        atomicAdd(&W2[e], W * (xb[a] + yb[b]));
                    }
            } else {
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const double Wip = sel3(i, wx0, wx1, wx2) * sel3(j, wy0, wy1, wy2);
                        const double d0 = (i - fx0) * h, d1 = (j - fx1) * h;
                        const double incV0 = Wip * (vx + (A00 * d0 + A01 * d1));
                        const double incV1 = Wip * (vy + (A10 * d0 + A11 * d1));
                        const int X = bx + i, Y = by + j;
                        if (X >= gx || Y >= gy) {
                            flag_error(k);
                            continue;
                        }
                        if ((unsigned)(X - ox) < (unsigned)kWX && (unsigned)(Y - oy) < (unsigned)kWY) {
                            const int e = (Y - oy) * kWS + (X - ox);
                            atomicAdd(&W0[e], Wip);
                            atomicAdd(&W1[e], incV0);
                            atomicAdd(&W2[e], incV1);
                        } else {
                            const int node = X + Y * gx;
                            unsafeAtomicAdd(&gO[node], Wip * m);
                            unsafeAtomicAdd(&gO[gpitch + node], incV0 * m);
                            unsafeAtomicAdd(&gO[2 * gpitch + node], incV1 * m);
                        }
                    }
            }
        }
        wsync();
        if constexpr (F32) __builtin_amdgcn_s_setprio(2);
        if (bcol < bw) {
            for (int ly = brow0; ly < bh; ly += bstep) {
                const int lx = bcol;
                const int e = ly * kWS + lx;
                // the three sums in one LDS round trip; the atomics address
                // the rows from their scalar bases with 32-bit offsets
                const double a = W0[e], a1 = W1[e], a2 = W2[e];
                if (a != 0) {
                    const int X = ox + lx, Y = oy + ly;
                    if (X < gx && Y < gy) {
                        const unsigned node = (unsigned)(X + Y * gx);
                        const double m = cp.mass;   // the window holds sums per unit mass
                        unsafeAtomicAdd(at(gO, node), a * m);
                        unsafeAtomicAdd(at(gO + gpitch, node), a1 * m);
                        unsafeAtomicAdd(at(gO + 2 * gpitch, node), a2 * m);
                    }
                }
            }
        }
    }
}

// P2G of a deferred point from the position, velocity and Bp its queue entry
// carries and the given Fe (rounded to the storage type first, as the pass
// continues from what it stores), straight to the grid with global atomics.
template <bool F32>
static __device__ void scatter_point(const MapEntry &e, const CaseParams &cp, const KParams &k,
                                     double F00, double F01, double F10, double F11,
                                     double *__restrict__ gridOut, size_t gpitch) {
    using V = typename Rows<F32>::V;
    constexpr double fe0 = F32 ? 1.0 : 0.0;
    const double px = e.x[0], py = e.x[1];
    const double vx = e.v[0], vy = e.v[1];
    const double B00 = e.B[0], B01 = e.B[1], B10 = e.B[2], B11 = e.B[3];
    if constexpr (F32) {
        F00 = fe0 + (double)(V)(F00 - fe0); F01 = (double)(V)F01;
        F10 = (double)(V)F10; F11 = fe0 + (double)(V)(F11 - fe0);
    }
    int bx, by;
    double fx0, fx1, wx0, wx1, wx2, wy0, wy1, wy2;
    if (!(px == px && py == py) || !stencil(px, cp.h_inv, bx, fx0, wx0, wx1, wx2) ||
        !stencil(py, cp.h_inv, by, fx1, wy0, wy1, wy2)) {
        flag_error(k);
        return;
    }
    {
        double t00, t01, t11;
        kirchhoff(F00, F01, F10, F11, cp.mu, cp.kappa, t00, t01, t11);
        const double m = cp.mass, h = cp.h;
        const double sc2 = cp.dt * cp.vol * cp.Dinv;
        const double A00 = m * B00 - sc2 * t00, A01 = m * B01 - sc2 * t01;
        const double A10 = m * B10 - sc2 * t01, A11 = m * B11 - sc2 * t11;
        const double mv0 = vx * m, mv1 = vy * m;
        double *__restrict__ gO = gridOut + cp.gridOffset;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const int X = bx + a, Y = by + b;
                if (X >= cp.gx || Y >= cp.gy) {
                    flag_error(k);
                    continue;
                }
                const double Wip = sel3(a, wx0, wx1, wx2) * sel3(b, wy0, wy1, wy2);
                const double d0 = (a - fx0) * h, d1 = (b - fx1) * h;
                const int node = X + Y * cp.gx;
                unsafeAtomicAdd(&gO[node], Wip * m);
                unsafeAtomicAdd(&gO[gpitch + node], Wip * (mv0 + (A00 * d0 + A01 * d1)));
                unsafeAtomicAdd(&gO[2 * gpitch + node], Wip * (mv1 + (A10 * d0 + A11 * d1)));
            }
    }
}

// The learned return map of the points a particle pass deferred: the MLP
// (one 16-row MFMA tile per wavefront and iteration, weights staged in
// shared memory) proposes the projected state of
// each queued point, the proposal is kept where its yield residual is <=
// tol and the exact FP64 map runs otherwise; then Fe and Jp_inv are stored
// and, when gridOut is given, the point's P2G contribution goes to the grid
// with global atomics (the pass left these points out of its P2G).  Resets
// the counter the next pass appends to.
template <bool F32>
__global__ __launch_bounds__(256) void k_batch_map(Rows<F32> S, const CaseParams *__restrict__ cases,
                                                  const MapEntry *__restrict__ q,
                                                  const unsigned int *__restrict__ qCount,
                                                  unsigned int *__restrict__ qCountNext,
                                                  double *__restrict__ gridOut, size_t gpitch,
                                                  int *__restrict__ err, LearnArgs la) {
    __shared__ mlpdev::InferScratch isc[4];
    __shared__ __attribute__((aligned(16))) unsigned short wsh[mlpdev::kParams];   // the weights, BF16
    using V = typename Rows<F32>::V;
    constexpr double fe0 = F32 ? 1.0 : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) qCountNext[0] = 0u;
    const unsigned int n = qCount[0];
    if (blockIdx.x * 64u >= n) return;   // no entry for this block (uniform)
    static_assert(mlpdev::kParams % 8 == 0, "weights are staged 8 at a time");
    for (int x = threadIdx.x; x < mlpdev::kParams / 8; x += 256)
        reinterpret_cast<uint4 *>(wsh)[x] = reinterpret_cast<const uint4 *>(la.wbf)[x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const size_t pp = S.pitch;
    V *__restrict__ R = S.rest;
    mlpdev::InferScratch &sc = isc[wv];
    // one 16-row tile per wavefront and iteration: lanes 0..15 own the entries
    for (unsigned int base = (blockIdx.x * 4 + wv) * 16u; base < n; base += gridDim.x * 64u) {
        const unsigned int i = base + lane;
        const bool have = lane < 16 && i < n;
        MapEntry e{};
        if (have) e = q[i];
        const CaseParams &cp = cases[have ? e.cs : 0];
        const KParams k = kp_of(cp, err);
        double F00 = have ? e.F[0] : 1.0, F01 = e.F[1], F10 = e.F[2], F11 = have ? e.F[3] : 1.0;
        double Jp = have ? e.Jp : 1.0;
        const Trial r = compute_pq_fast(F00, F01, F10, F11, cp.mu, cp.kappa);
        double qlim = 0;
        const int br = map_branch(k, r, Jp, qlim);
        const bool need = have && br != kBrInside && r.svd.s1 > 0;
        double pred[3] = {0, 0, 0};
        if (__ballot(need)) {
            for (int x = lane; x < 16 * mlpdev::kIn; x += 64) sc.x[x / mlpdev::kIn][x % mlpdev::kIn] = 0;
            mlpdev::wave_fence();
            if (need) {
                double f[5];
                map_features(r, Jp, br, f);
                for (int t = 0; t < 5; ++t)
                    sc.x[lane][t] = mlpdev::to_bf16((float)((f[t] - la.norm.in_mean[t]) / la.norm.in_scale[t]));
            }
            mlpdev::wave_fence();
            mlpdev::infer16(sc, wsh, la.pf);
            if (need)
                for (int t = 0; t < 3; ++t)
                    pred[t] = (double)sc.y[lane][t] * la.norm.out_scale[t] + la.norm.out_mean[t];
            mlpdev::wave_fence();
        }
        bool accepted = false;
        if (need) {
            const double s0 = r.svd.s0 * exp(pred[0]);   // log s0 + pred0, exponentiated
            const double s1 = r.svd.s1 * exp(pred[1]);
            // the shear branch leaves Jp_inv unchanged
            const double Jpn = (br == kBrTension) ? Jp * exp(pred[2]) : Jp;
            if (yield_residual(k, r, Jp, br, qlim, s0, s1, Jpn) <= la.tol) {
                usvt(r.svd, s0, s1, F00, F01, F10, F11);
                Jp = Jpn;
                accepted = true;
            }
        }
        if (have && !accepted) dp_return(k, r, F00, F01, F10, F11, Jp);
        const unsigned long long acc = __ballot(accepted), fb = __ballot(have && !accepted);
        const unsigned long long ten = __ballot(have && br == kBrTension);
        if (lane == 0) {
            if (acc) atomicAdd(&la.stats[0], (unsigned long long)__popcll(acc));
            if (fb) atomicAdd(&la.stats[1], (unsigned long long)__popcll(fb));
            if (acc & ten) atomicAdd(&la.stats[3], (unsigned long long)__popcll(acc & ten));
            if (fb & ten) atomicAdd(&la.stats[4], (unsigned long long)__popcll(fb & ten));
            if (acc & ~ten) atomicAdd(&la.stats[5], (unsigned long long)__popcll(acc & ~ten));
            if (fb & ~ten) atomicAdd(&la.stats[6], (unsigned long long)__popcll(fb & ~ten));
        }
        if (have) {
            const int pt = e.pt;
            if constexpr (F32) {
                fe_rec(R, pp)[pt] = fe4{(float)(F00 - fe0), (float)F01, (float)F10, (float)(F11 - fe0)};
            } else {
                R[pp * 2 + pt] = (V)(F00 - fe0);
                R[pp * 3 + pt] = (V)F01;
                R[pp * 4 + pt] = (V)F10;
                R[pp * 5 + pt] = (V)(F11 - fe0);
            }
            if (Jp != e.Jp) R[pp * 10 + pt] = (V)(Jp - fe0);
            if (gridOut) scatter_point<F32>(e, cp, k, F00, F01, F10, F11, gridOut, gpitch);
        }
    }
}

// Grid update over all cases: block b handles nodes [nodeBlock[b].y, +256)
// of case nodeBlock[b].x.  The single-case update (gpu_implementation5.cu:
// 215-290) with three FAST forms: 1/m from rcp_nr, the speed limit tested on
// the squared speed, the node's row from a reciprocal of the grid width.
// The substep's total indenter force of each case goes to hist[case*2 +
// comp].  The velocities go to rows 1-2 of the grid, or, when vel32 is given
// (FP32 point storage), to vel32 as FP32 pairs, zero for the nodes without
// mass.  A node's mass and momentum are loaded together, so that a block
// waits for memory once.
// Per-block flags skip the stores that would write zeros over zeros: a
// block of zeroGrid is cleared only if it took mass in its last P2G
// (zeroHasMass, recorded when that grid was last updated), and a block of
// vel32 is rewritten at its massless nodes only if it last held a velocity
// (velNonzero).  The host sets every flag when a run starts without a
// P2G left by the previous one.
__global__ __launch_bounds__(256) void k_batch_update(const CaseParams *__restrict__ cases,
                                                     const int2 *__restrict__ nodeBlock,
                                                     double *__restrict__ grid, double *__restrict__ zeroGrid,
                                                     size_t gpitch, const double *__restrict__ tTab, int step,
                                                     double *__restrict__ hist, float2 *__restrict__ vel32,
                                                     int *__restrict__ gridHasMass,
                                                     const int *__restrict__ zeroHasMass,
                                                     int *__restrict__ velNonzero, double *__restrict__ bins) {
    __shared__ double sf[2];
    __shared__ int any;
    const int2 nb = nodeBlock[blockIdx.x];
    const CaseParams &cp = cases[nb.x];
    const int local = nb.y + (int)threadIdx.x;
    const bool node = local < cp.gridTotal;
    const size_t idx = (size_t)cp.gridOffset + (node ? local : 0);
    double mass = 0, p0 = 0, p1 = 0;
    if (node) {
        mass = grid[idx];
        p0 = grid[gpitch + idx];
        p1 = grid[2 * gpitch + idx];
    }
    if (threadIdx.x == 0) {
        sf[0] = 0;
        sf[1] = 0;
        any = 0;
    }
    // a block of zeroGrid that took no mass in its last P2G is still zero,
    // and a block of vel32 last written with no mass holds zeros
    const bool zeroIt = zeroGrid && zeroHasMass[blockIdx.x] != 0;
    const bool velDirty = vel32 && velNonzero[blockIdx.x] != 0;
    const bool blockMass = __syncthreads_or(mass != 0) != 0;   // also orders sf/any
    if (threadIdx.x == 0) {
        gridHasMass[blockIdx.x] = blockMass ? 1 : 0;
        if (vel32) velNonzero[blockIdx.x] = blockMass ? 1 : 0;
    }
    if (node) {
        if (zeroIt) {
            zeroGrid[idx] = 0;
            zeroGrid[gpitch + idx] = 0;
            zeroGrid[2 * gpitch + idx] = 0;
        }
        if (velDirty && mass == 0) vel32[idx] = make_float2(0.f, 0.f);
        if (mass != 0) {
            const double dt = cp.dt, h = cp.h, V = cp.indV;
            const double t = tTab[step];
            const double ind_x = cp.x0 + t * V, ind_y = cp.y0;
            const double rm = rcp_nr(mass);
            double v0 = p0 * rm, v1 = p1 * rm;
            v1 = v1 - dt * cp.gravity;
            const double vmax = 0.5 * h / dt;
            const double vsq = v0 * v0 + v1 * v1;
            if (vsq > vmax * vmax) {
                const double s = vmax / sqrt(vsq);
                v0 = v0 * s;
                v1 = v1 * s;
            }
            // row and column: (local + 0.5) / gx lies >= 0.5 / gx >= 2^-17
            // from an integer and is computed to within 2^-23, so its floor
            // is the row
            const int iy = (int)(((double)local + 0.5) * cp.gxInv);
            const int ix = local - iy * cp.gx;
            double n0 = ix * h - ind_x, n1 = iy * h - ind_y;
            if (n0 * n0 + n1 * n1 < cp.indRsq) {
                const double r0 = v0 - V, r1 = v1;
                const double z = n0 * n0 + n1 * n1;
                if (z > 0) {
                    const double nz = sqrt(z);
                    n0 = n0 / nz;
                    n1 = n1 / nz;
                }
                const double vnr = r0 * n0 + r1 * n1;
                if (vnr < 0) {
                    const double q0 = v0, q1 = v1;
                    v0 = V + (r0 - n0 * vnr);
                    v1 = r1 - n1 * vnr;
                    const double f0 = (q0 - v0) * mass / dt, f1 = (q1 - v1) * mass / dt;
                    atomicAdd(&sf[0], f0);
                    atomicAdd(&sf[1], f1);
                    any = 1;
                    // the force's angular bin (gpu_implementation5.cu:262-269:
                    // atan2 of (x, y) of the unit normal, then + pi, scaled)
                    double angle = atan2(n0, n1);
                    angle += kPi;
                    angle *= cp.nbins / (2 * kPi);
                    int kb = (int)angle;
                    kb = max(kb, 0);
                    kb = min(kb, cp.nbins - 1);
                    unsafeAtomicAdd(&bins[cp.binOffset + 2 * kb], f0);
                    unsafeAtomicAdd(&bins[cp.binOffset + 2 * kb + 1], f1);
                }
            }
            if (iy <= 2) {
                v0 = 0;
                v1 = 0;
            } else if (iy >= cp.gy - 3 && v1 > 0) {
                v1 = 0;
            }
            if (ix <= 2 && v0 < 0)
                v0 = 0;
            else if (ix >= cp.gx - 3 && v0 > 0)
                v0 = 0;
            if (vel32) {
                vel32[idx] = make_float2((float)v0, (float)v1);
            } else {
                grid[gpitch + idx] = v0;
                grid[2 * gpitch + idx] = v1;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && any) {
        unsafeAtomicAdd(&hist[2 * nb.x], sf[0]);
        unsafeAtomicAdd(&hist[2 * nb.x + 1], sf[1]);
    }
}

// Cell key of every point of the batch, offset by its case's first node so
// that a sort keeps the cases contiguous and in order; vals = slot index.
template <class PT>
__global__ void k_batch_keys(const PT *__restrict__ pos, size_t pitch, int nPts,
                             const CaseParams *__restrict__ cases, int nCases, int *__restrict__ keys,
                             int *__restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPts) return;
    int lo = 0, hi = nCases - 1;   // the last case with ptOffset <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cases[mid].ptOffset <= i) lo = mid; else hi = mid - 1;
    }
    const CaseParams &cp = cases[lo];
    double x, y;
    pos_load<false>(pos, pitch, i, cp, x, y);
    int key = cp.gridTotal - 1;
    if (x == x && y == y) {
        const double bx = fmin(fmax(x * cp.h_inv - 0.5, 0.0), (double)(cp.gx - 1));
        const double by = fmin(fmax(y * cp.h_inv - 0.5, 0.0), (double)(cp.gy - 1));
        key = band_key((int)bx, (int)by, cp.gx, cp.gy);
    }
    keys[i] = cp.gridOffset + key;
    vals[i] = i;
}

// The case-local base cell of every slot, packed as column | row << 16
// (finalize keeps every grid within 65536 columns and rows): what the chunk
// table is built from, without the band key's divisions on the host.
template <class PT>
__global__ void k_batch_cells(const PT *__restrict__ pos, size_t pitch, int nPts,
                              const CaseParams *__restrict__ cases, int nCases, unsigned *__restrict__ cells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPts) return;
    int lo = 0, hi = nCases - 1;   // the last case with ptOffset <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cases[mid].ptOffset <= i) lo = mid; else hi = mid - 1;
    }
    const CaseParams &cp = cases[lo];
    double x, y;
    pos_load<false>(pos, pitch, i, cp, x, y);
    // the cell k_batch_keys keys the slot by (a NaN position: the last cell)
    unsigned cx = (unsigned)(cp.gx - 1), cy = (unsigned)(cp.gy - 1);
    if (x == x && y == y) {
        cx = (unsigned)fmin(fmax(x * cp.h_inv - 0.5, 0.0), (double)(cp.gx - 1));
        cy = (unsigned)fmin(fmax(y * cp.h_inv - 0.5, 0.0), (double)(cp.gy - 1));
    }
    cells[i] = cx | cy << 16;
}

// Point state permuted: slot k of the destination <- slot perm[k] of the source.
template <bool F32>
__global__ void k_batch_permute(Rows<F32> src, Rows<F32> dst, const int *__restrict__ srcOrig,
                                int *__restrict__ dstOrig, const int *__restrict__ perm, int nPts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nPts) return;
    const int j = perm[k];
    const size_t p = src.pitch;
    if constexpr (F32)
        reinterpret_cast<prec4 *>(dst.pos)[k] = reinterpret_cast<const prec4 *>(src.pos)[j];
    else
        for (int r = 0; r < Rows<F32>::kPosRows; ++r) dst.pos[r * p + k] = src.pos[r * p + j];
    if constexpr (F32) {
        for (int r = 0; r < 11; ++r)
            if (r < 2 || r > 5) dst.rest[r * p + k] = src.rest[r * p + j];
        fe_rec(dst.rest, p)[k] = fe_rec(static_cast<const float *>(src.rest), p)[j];
    } else {
        for (int r = 0; r < 11; ++r) dst.rest[r * p + k] = src.rest[r * p + j];
    }
    dst.crushed[k] = src.crushed[j];
    if (src.grain) dst.grain[k] = src.grain[j];
    dstOrig[k] = srcOrig[j];
}

}  // namespace batch
}  // namespace icempm

using namespace icempm;
using namespace icempm::batch;

struct icempm_batch {
    int precision = ICEMPM_BATCH_F64;
    hipStream_t stream = nullptr;
    // The cases in kChains groups of about equal point counts run their
    // substep loops as independent chains (group g > 0 on side[g - 1]), so
    // that one group's grid update, bound by memory, overlaps another's
    // particle pass, bound by issue.  groupCase: each group's first case;
    // grpChunk / grpBlock: the groups' ranges of the chunk table and of the
    // node blocks; chains: the number of groups (1 when there is one case).
    static constexpr int kChains = ICEMPM_BATCH_CHAINS;
    hipStream_t side[kChains - 1] = {};
    hipEvent_t evFork = nullptr;
    hipEvent_t evJoin[kChains - 1] = {};
    int chains = 1;
    int groupCase[kChains + 1] = {};
    int grpChunk[kChains + 1] = {};
    int grpBlock[kChains + 1] = {};
    // the launch target of the next particle pass: its stream and chunk range
    hipStream_t lstream = nullptr;
    int lchunk0 = 0, lchunks = 0;
    std::string lastError;
    bool finalized = false;

    struct HostCase {
        CaseParams cp;
        std::vector<double> pos;        // 2 x n
        std::vector<double> rest;       // 11 x n
        std::vector<unsigned char> crushed;
        std::vector<short> grain;
        std::vector<int> keys;
    };
    std::vector<HostCase> hc;

    // device
    void *pos = nullptr;          // Rows<F32>::kPosRows rows of Rows<F32>::P
    void *rest = nullptr;
    unsigned char *crushed = nullptr;
    short *grain = nullptr;       // per-point grain ids (when a case has GrainVariability != 0)
    double *bins = nullptr;       // indenter force bins of every case, summed over a run
    int binsTotal = 0;            // doubles
    size_t pitch = 0;
    int nPts = 0;
    double *gridA = nullptr, *gridB = nullptr;
    float2 *vel32 = nullptr;   // FP32 storage: the grid velocities the gathers read
    int *blkFlags = nullptr;   // per node block: grid A took mass, grid B took mass, vel32 nonzero
    size_t gpitch = 0;
    long nNodes = 0;
    CaseParams *cases = nullptr;
    int4 *chunks = nullptr;   // {first point, end, case, 0} per chunk
    int nChunks = 0;
    size_t chunkCapDev = 0;
    int2 *nodeBlock = nullptr;
    int nNodeBlocks = 0;
    int *err = nullptr;
    double *tTab = nullptr;
    double *hist = nullptr;
    int cap = 0;               // substeps the buffers hold
    int lastRun = 0;

    double t = 0;
    int step = 0;

    // return map of crushed points
    int rmap = kMapExact;
    unsigned short *mlpW = nullptr;   // MLP parameters as BF16 bits
    float *mlpP = nullptr;            // and as FP32
    mlpdev::Norm norm{};
    double tol = 1e-4;
    float *hx = nullptr, *ht = nullptr;
    unsigned long long *hcount = nullptr;
    int hcap = 0;
    unsigned long long *stats = nullptr;   // kMapStats counters
    int mapBranches = kMapTension | kMapShear;   // branches the network proposes for
    double mapMinAccept = 0.5;         // autotune: a branch accepted less often is switched off
    unsigned long long tuneBranchStats[4] = {};   // the last autotune's per-branch accepted / fell back
    MapEntry *mapQ = nullptr;          // deferred points (learned map), nPts entries
    unsigned int *mapCount = nullptr;  // two counters, used alternately

    // executable graphs of the substep loop, two per start (of different
    // step counts, the least recently used replaced): fresh (row 0), or the
    // first substep's P2G already in grid A (row 1) or grid B (row 2); each
    // with the step count, time table and history rows it was captured
    // with, and the grid that holds the next P2G after it
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int steps = 0;
        const double *T = nullptr;
        double *H = nullptr;
        int end = -1;
        unsigned long long used = 0;
    };
    GraphSlot graphs[3][2];
    unsigned long long graphTick = 0;
    // the grid (0: A, 1: B) that holds the P2G of the next substep, made
    // from the current points by the last pass of the previous run; -1 when
    // there is none, or the points have changed since (restore, re-sort)
    int pending = -1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float lastMs = 0;

    // tuning: points per chunk (= threads per workgroup of the particle pass)
    // and the re-sort cadence in substeps (0: never)
    int chunkPoints = 256;
    int resortEvery = 0;
    std::vector<int> hcstart, hccase;       // the chunk table on the host
    int *orig = nullptr;                    // batch index of the point in each slot
    // re-sort scratch
    int *sKeys = nullptr, *sVals = nullptr, *sKeys2 = nullptr, *sVals2 = nullptr, *sTmp = nullptr;
    // the slots' base cells for a new chunk table: on the device, and read
    // back into pinned host memory
    unsigned *dcells = nullptr, *hcells = nullptr;
    void *tpos = nullptr;
    void *trest = nullptr;
    unsigned char *tcrushed = nullptr;
    short *tgrain = nullptr;
    int *torig = nullptr;
    // checkpoint of the state
    bool haveCheckpoint = false;
    void *kpos = nullptr;
    void *krest = nullptr;
    unsigned char *kcrushed = nullptr;
    short *kgrain = nullptr;
    int *korig = nullptr;
    double kt = 0;
    int kstep = 0;
    std::vector<int> kcstart, kccase;
    int kChunkPoints = 256;
    double runMs = 0;   // device time of the last run's segments + re-sorts

    // device timing of the fused particle pass (stream launches only)
    int timingStride = 0;
    std::vector<hipEvent_t> tev;   // pairs
    hipEvent_t evAfter = nullptr;  // while set: recorded after the next particle kernel
    bool storeVB = false;          // while set: the next particle kernel stores v and Bp
    size_t tevNext = 0;
    double timedMs = 0;
    int timedLaunches = 0;
};

namespace {

int bfail(icempm_batch *b, int code, const std::string &m) {
    if (b) b->lastError = m;
    return code;
}

#define BCHK(b, call, tag)                                                                   \
    do {                                                                                     \
        hipError_t _e = (call);                                                              \
        if (_e != hipSuccess) return bfail(b, ICEMPM_E_HIP, std::string(tag) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

constexpr int kMapBlocks = 256;

template <bool F32, bool G2P, bool P2G, int RMAP>
void launch_one(icempm_batch *b, unsigned nb, const void *gV, double *gO, const LearnArgs &la,
                const MapQueue &mq) {
    Rows<F32> S{static_cast<typename Rows<F32>::P *>(b->pos), reinterpret_cast<typename Rows<F32>::V *>(b->rest),
                b->crushed, b->grain, b->pitch};
    hipLaunchKernelGGL((k_batch_particles<F32, G2P, P2G, RMAP>), dim3(nb), dim3(256), 0, b->lstream,
                       b->chunks + b->lchunk0, b->lchunks, b->cases, S.pos, S.rest, S.crushed, S.pitch,
                       static_cast<const typename Rows<F32>::G *>(gV), gO, b->gpitch,
                       b->err, b->grain, la, mq, b->storeVB ? 1 : 0);
    if (b->evAfter) (void)hipEventRecord(b->evAfter, b->lstream);
    if (G2P && RMAP == kMapLearned && la.branches != 0) {
        // the deferred points: map, store, and their P2G into the pass's output grid
        hipLaunchKernelGGL((k_batch_map<F32>), dim3(kMapBlocks), dim3(256), 0, b->lstream, S, b->cases, mq.q,
                           mq.count, mq.count == b->mapCount ? b->mapCount + 1 : b->mapCount, gO, b->gpitch,
                           b->err, la);
    }
}

template <bool F32, int RMAP>
void launch_map(icempm_batch *b, unsigned nb, bool g2p, bool p2g, const void *gV, double *gO,
                const LearnArgs &la, const MapQueue &mq) {
    if (g2p && p2g)
        launch_one<F32, true, true, RMAP>(b, nb, gV, gO, la, mq);
    else if (g2p)
        launch_one<F32, true, false, RMAP>(b, nb, gV, gO, la, mq);
    else
        launch_one<F32, false, true, kMapExact>(b, nb, gV, gO, la, mq);
}

template <bool F32>
void launch_prec(icempm_batch *b, unsigned nb, bool g2p, bool p2g, const void *gV, double *gO,
                 const LearnArgs &la, const MapQueue &mq) {
    // the learned map with every branch switched off is the exact map: its
    // pass, without the queue logic and the learned map's counters
    if (b->rmap == kMapHarvest)
        launch_map<F32, kMapHarvest>(b, nb, g2p, p2g, gV, gO, la, mq);
    else if (b->rmap == kMapLearned && b->mapBranches != 0)
        launch_map<F32, kMapLearned>(b, nb, g2p, p2g, gV, gO, la, mq);
    else
        launch_map<F32, kMapExact>(b, nb, g2p, p2g, gV, gO, la, mq);
}

// One particle pass (and, with the learned map, the map kernel after it);
// parity selects the queue counter the pass appends to.
// gV: the grid whose rows 1-2 hold the velocities (FP64 storage) or vel32
hipError_t launch_pass(icempm_batch *b, bool g2p, bool p2g, const void *gV, double *gO, int parity,
                       hipStream_t st, int chunk0, int chunks) {
    const unsigned nb = (unsigned)(((chunks + 7) / 8) * 8);
    if (chunks <= 0) return hipSuccess;
    b->lstream = st;
    b->lchunk0 = chunk0;
    b->lchunks = chunks;
    LearnArgs la{};
    la.wbf = b->mlpW;
    la.pf = b->mlpP;
    la.norm = b->norm;
    la.tol = b->tol;
    la.hx = b->hx;
    la.ht = b->ht;
    la.hcount = b->hcount;
    la.hcap = b->hcap;
    la.stats = b->stats;
    la.branches = b->mapBranches;
    MapQueue mq{b->mapQ, b->mapCount ? b->mapCount + (parity & 1) : nullptr, b->nPts};
    if (b->precision == ICEMPM_BATCH_F32)
        launch_prec<true>(b, nb, g2p, p2g, gV, gO, la, mq);
    else
        launch_prec<false>(b, nb, g2p, p2g, gV, gO, la, mq);
    return hipGetLastError();
}

// Chunk table from the case-local base cells of the points in slot order
// (column | row << 16): chunks of <= chunkPoints consecutive points of one
// case whose base cells span at most two cell rows and 126 columns, so that
// a wavefront's window (4 node rows) holds all its stencils, cut greedily.
// Points in band order give chunks of one band each, until they move.  The
// cases are independent, so they are cut on several host threads.
void chunk_table(const icempm_batch *b, const unsigned *cells, std::vector<int> &cstart,
                 std::vector<int> &ccase) {
    const size_t nc = b->hc.size();
    const int cap = b->chunkPoints;
    std::vector<std::vector<int>> starts(nc);
    auto cut = [&](size_t ci) {
        const CaseParams &cp = b->hc[ci].cp;
        const int n = cp.nPts, o = cp.ptOffset;
        std::vector<int> &st = starts[ci];
        if (n <= 0) return;
        st.reserve((size_t)n / (size_t)cap + 16);
        st.push_back(o);
        int cs = 0;
        int c0 = (int)(cells[o] & 0xFFFFu), r0 = (int)(cells[o] >> 16);
        int c1 = c0, r1 = r0;
        for (int i = 1; i < n; ++i) {
            const unsigned ki = cells[(size_t)o + i];
            const int cc = (int)(ki & 0xFFFFu), r = (int)(ki >> 16);
            const int nr0 = std::min(r0, r), nr1 = std::max(r1, r), nc0 = std::min(c0, cc), nc1 = std::max(c1, cc);
            if ((i - cs >= cap) || nr1 - nr0 > 1 || nc1 - nc0 > 125) {
                st.push_back(o + i);
                cs = i;
                r0 = r1 = r;
                c0 = c1 = cc;
            } else {
                r0 = nr0; r1 = nr1; c0 = nc0; c1 = nc1;
            }
        }
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nt = std::min<size_t>(nc, std::min(16u, std::max(1u, hw)));
    if (nt <= 1 || b->nPts < (1 << 16)) {
        for (size_t ci = 0; ci < nc; ++ci) cut(ci);
    } else {
        // the calling thread works too; a helper that cannot be started
        // leaves its share to the others, and an exception (allocation)
        // in any of them is rethrown here, as the serial loop would throw
        // (the entry point's host_guard turns it into ICEMPM_E_HOST)
        std::atomic<size_t> next{0};
        std::exception_ptr failure;
        std::mutex failureLock;
        auto work = [&] {
            try {
                for (size_t ci; (ci = next.fetch_add(1)) < nc;) cut(ci);
            } catch (...) {
                std::lock_guard<std::mutex> g(failureLock);
                if (!failure) failure = std::current_exception();
            }
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t) {
            try {
                pool.emplace_back(work);
            } catch (const std::system_error &) {
                break;
            }
        }
        work();
        for (auto &t : pool) t.join();
        if (failure) std::rethrow_exception(failure);
    }
    cstart.clear();
    ccase.clear();
    for (size_t ci = 0; ci < nc; ++ci)
        for (int st : starts[ci]) {
            cstart.push_back(st);
            ccase.push_back((int)ci);
        }
    cstart.push_back(b->nPts);
}

void drop_graph(icempm_batch *b) {
    for (auto &row : b->graphs)
        for (auto &g : row) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g = icempm_batch::GraphSlot{};
        }
}

// the host chunk table to the device (grows the device array when needed)
int upload_chunks(icempm_batch *b) {
    const size_t n = b->hccase.size();
    std::vector<int4> t(std::max<size_t>(n, 1));
    for (size_t c = 0; c < n; ++c) t[c] = make_int4(b->hcstart[c], b->hcstart[c + 1], b->hccase[c], 0);
    const int4 *before = b->chunks;
    const int nBefore = b->nChunks;
    if (t.size() > b->chunkCapDev) {
        dfree(b->chunks);
        b->chunkCapDev = 0;
        BCHK(b, hipMalloc(&b->chunks, t.size() * sizeof(int4)), "chunk table");
        b->chunkCapDev = t.size();
    }
    BCHK(b, hipMemcpy(b->chunks, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice), "chunk table");
    b->nChunks = (int)n;
    // each group's first chunk
    bool splitMoved = false;
    for (int g = 0; g <= b->chains; ++g) {
        const int c = g == b->chains ? (int)n
                                     : (int)(std::lower_bound(b->hccase.begin(), b->hccase.end(), b->groupCase[g]) -
                                             b->hccase.begin());
        splitMoved = splitMoved || c != b->grpChunk[g];
        b->grpChunk[g] = c;
    }
    // a captured loop holds the table's address, the chunk count and the
    // groups' ranges (the grid sizes of its launches); the table's contents
    // are read at run time
    if (b->chunks != before || b->nChunks != nBefore || splitMoved) drop_graph(b);
    return ICEMPM_OK;
}

// the sort keys (batch-wide band keys) and slot indices of the current
// slots in sKeys / sVals
int sort_keys(icempm_batch *b) {
    const int n = b->nPts;
    if (!b->sKeys) {
        BCHK(b, hipMalloc(&b->sKeys, (size_t)n * sizeof(int)), "keys");
        BCHK(b, hipMalloc(&b->sVals, (size_t)n * sizeof(int)), "keys");
    }
    if (b->precision == ICEMPM_BATCH_F32)
        hipLaunchKernelGGL(k_batch_keys<unsigned int>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                           static_cast<const unsigned int *>(b->pos), b->pitch, n, b->cases, (int)b->hc.size(),
                           b->sKeys, b->sVals);
    else
        hipLaunchKernelGGL(k_batch_keys<double>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                           static_cast<const double *>(b->pos), b->pitch, n, b->cases, (int)b->hc.size(), b->sKeys,
                           b->sVals);
    BCHK(b, hipGetLastError(), "keys");
    return ICEMPM_OK;
}

// the base cells of the current slots, in b->hcells on the host
int current_cells(icempm_batch *b) {
    const int n = b->nPts;
    if (!b->dcells) {
        BCHK(b, hipMalloc(&b->dcells, (size_t)n * sizeof(unsigned)), "cells");
        BCHK(b, hipHostMalloc(&b->hcells, (size_t)n * sizeof(unsigned), hipHostMallocDefault), "cells");
    }
    if (b->precision == ICEMPM_BATCH_F32)
        hipLaunchKernelGGL(k_batch_cells<unsigned int>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                           static_cast<const unsigned int *>(b->pos), b->pitch, n, b->cases, (int)b->hc.size(),
                           b->dcells);
    else
        hipLaunchKernelGGL(k_batch_cells<double>, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                           static_cast<const double *>(b->pos), b->pitch, n, b->cases, (int)b->hc.size(), b->dcells);
    BCHK(b, hipGetLastError(), "cells");
    BCHK(b, hipMemcpyAsync(b->hcells, b->dcells, (size_t)n * sizeof(unsigned), hipMemcpyDeviceToHost, b->stream),
         "cells");
    BCHK(b, hipStreamSynchronize(b->stream), "cells");
    return ICEMPM_OK;
}

size_t pos_bytes(const icempm_batch *b) {
    return b->precision == ICEMPM_BATCH_F32 ? Rows<true>::kPosRows * b->pitch * sizeof(unsigned int)
                                             : Rows<false>::kPosRows * b->pitch * sizeof(double);
}

size_t rest_bytes(const icempm_batch *b) {
    return 11 * b->pitch * (b->precision == ICEMPM_BATCH_F32 ? sizeof(float) : sizeof(double));
}

// Element of rest value r (v0, v1, Fe00, Fe01, Fe10, Fe11, Bp00..Bp11, Jp_inv)
// of a slot: a row each, except FP32 storage's Fe, one 4-float record per
// slot in the space of rows 2-5.
size_t rest_elem(const icempm_batch *b, int r, size_t slot) {
    if (b->precision == ICEMPM_BATCH_F32 && r >= 2 && r <= 5) return 2 * b->pitch + 4 * slot + (size_t)(r - 2);
    return (size_t)r * b->pitch + slot;
}

// Stable re-sort of every case's points by their current cell, then a new
// chunk table.  The slot -> batch-index map (orig) follows the points, so
// reads return them in their original order.
int resort(icempm_batch *b) {
    const int n = b->nPts;
    if (n <= 1) return ICEMPM_OK;
    int rc = sort_keys(b);
    if (rc) return rc;
    if (!b->sKeys2) {
        BCHK(b, hipMalloc(&b->sKeys2, (size_t)n * sizeof(int)), "re-sort");
        BCHK(b, hipMalloc(&b->sVals2, (size_t)n * sizeof(int)), "re-sort");
        BCHK(b, hipMalloc(&b->sTmp, radix_tmp_ints((size_t)n) * sizeof(int)), "re-sort");
        BCHK(b, hipMalloc(&b->tpos, pos_bytes(b)), "re-sort");
        BCHK(b, hipMalloc(&b->trest, rest_bytes(b)), "re-sort");
        BCHK(b, hipMalloc(&b->tcrushed, b->pitch), "re-sort");
        BCHK(b, hipMalloc(&b->torig, b->pitch * sizeof(int)), "re-sort");
        if (b->grain) BCHK(b, hipMalloc(&b->tgrain, b->pitch * sizeof(short)), "re-sort");
    }
    BCHK(b, radix_sort_pairs(b->sKeys, b->sVals, b->sKeys2, b->sVals2, (size_t)n, b->sTmp, b->stream), "re-sort");
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (b->precision == ICEMPM_BATCH_F32) {
        Rows<true> a{static_cast<unsigned int *>(b->pos), reinterpret_cast<float *>(b->rest), b->crushed, b->grain,
                     b->pitch};
        Rows<true> t{static_cast<unsigned int *>(b->tpos), reinterpret_cast<float *>(b->trest), b->tcrushed,
                     b->tgrain, b->pitch};
        hipLaunchKernelGGL(k_batch_permute<true>, dim3(nb), dim3(256), 0, b->stream, a, t, b->orig, b->torig, b->sVals, n);
    } else {
        Rows<false> a{static_cast<double *>(b->pos), reinterpret_cast<double *>(b->rest), b->crushed, b->grain,
                      b->pitch};
        Rows<false> t{static_cast<double *>(b->tpos), reinterpret_cast<double *>(b->trest), b->tcrushed,
                      b->tgrain, b->pitch};
        hipLaunchKernelGGL(k_batch_permute<false>, dim3(nb), dim3(256), 0, b->stream, a, t, b->orig, b->torig, b->sVals, n);
    }
    BCHK(b, hipGetLastError(), "re-sort");
    std::swap(b->pos, b->tpos);
    std::swap(b->rest, b->trest);
    std::swap(b->crushed, b->tcrushed);
    std::swap(b->grain, b->tgrain);
    std::swap(b->orig, b->torig);
    drop_graph(b);   // the captured launches hold the old row addresses
    b->pending = -1;   // a P2G in a grid scattered the points from their old slots
    // the cells of the sorted slots give the new chunk table
    if ((rc = current_cells(b))) return rc;
    chunk_table(b, b->hcells, b->hcstart, b->hccase);
    return upload_chunks(b);
}

// the chunk table for the current slots and chunkPoints
int rebuild_chunks(icempm_batch *b) {
    int rc = current_cells(b);
    if (rc) return rc;
    chunk_table(b, b->hcells, b->hcstart, b->hccase);
    return upload_chunks(b);
}

// The substep loop of n substeps starting from grid A, as stream work.
// The substep loop: n times the grid update and the fused pass (G2P of
// this substep, P2G of the next into the other grid).  A fresh start (start
// < 0) clears grid A and scatters the points into it first; start = 0 or 1
// takes the P2G the previous run's last pass left in grid A or B, with the
// block flags that run left.  *end receives the grid that then holds the
// next substep's P2G.
hipError_t enqueue_loop(icempm_batch *b, int n, const double *tTab, double *hist, int start, int *end) {
    hipError_t e;
    const size_t nblk = (size_t)b->nNodeBlocks;
    double *A0 = b->gridA, *B0 = b->gridB;
    int *fA0 = b->blkFlags, *fB0 = b->blkFlags + nblk, *fV = b->blkFlags + 2 * nblk;
    if (start < 0 && (e = launch_fill_f64(b->gridA, b->gpitch * 3, 0.0, b->stream)) != hipSuccess) return e;
    if ((e = launch_fill_f64(hist, 2 * (size_t)n * b->hc.size(), 0.0, b->stream)) != hipSuccess) return e;
    if (b->mapCount &&
        (e = launch_fill_i32(reinterpret_cast<int *>(b->mapCount), 2, 0, b->stream)) != hipSuccess)
        return e;
    int cur = 0;   // which grid A is
    if (start < 0) {
        if ((e = launch_pass(b, false, true, b->gridA, b->gridA, 0, b->stream, 0, b->nChunks)) != hipSuccess)
            return e;
        // the grids' and vel32's contents are unknown at a fresh start
        if ((e = launch_fill_i32(b->blkFlags, 3 * nblk, 1, b->stream)) != hipSuccess) return e;
    } else if (start == 1) {
        std::swap(A0, B0);
        std::swap(fA0, fB0);
        cur = 1;
    }
    // The groups of cases as independent chains, when there are several,
    // the map kernel (one queue for all cases) is not in the loop and no
    // pass is being timed: the side chains start after everything above,
    // and the main stream waits for them at the end.
    const bool mapKernel = b->rmap == kMapLearned && b->mapBranches != 0;
    const int groups = (!mapKernel && b->timingStride <= 0) ? b->chains : 1;
    if (groups > 1) {
        if ((e = hipEventRecord(b->evFork, b->stream)) != hipSuccess) return e;
        for (int g = 1; g < groups; ++g)
            if ((e = hipStreamWaitEvent(b->side[g - 1], b->evFork, 0)) != hipSuccess) return e;
    }
    const size_t nc = b->hc.size();
    for (int g = 0; g < groups; ++g) {
        hipStream_t st = g == 0 ? b->stream : b->side[g - 1];
        const int c0 = groups > 1 ? b->grpChunk[g] : 0;
        const int nch = groups > 1 ? b->grpChunk[g + 1] - c0 : b->nChunks;
        const int k0 = groups > 1 ? b->grpBlock[g] : 0;
        const int nkb = groups > 1 ? b->grpBlock[g + 1] - k0 : b->nNodeBlocks;
        double *A = A0, *B = B0;
        int *fA = fA0, *fB = fB0;
        for (int i = 0; i < n; ++i) {
            if (nkb > 0) {
                hipLaunchKernelGGL(k_batch_update, dim3(nkb), dim3(256), 0, st, b->cases, b->nodeBlock + k0, A, B,
                                   b->gpitch, tTab, i, hist + 2 * nc * (size_t)i, b->vel32, fA + k0, fB + k0,
                                   fV + k0, b->bins);
                if ((e = hipGetLastError()) != hipSuccess) return e;
            }
            const void *gV = b->vel32 ? static_cast<const void *>(b->vel32) : static_cast<const void *>(A);
            const bool timeIt = b->timingStride > 0 && i % b->timingStride == 0;
            hipEvent_t t0 = nullptr, t1 = nullptr;
            if (timeIt) {
                if (b->tevNext + 2 > b->tev.size()) {
                    for (int q = 0; q < 2; ++q) {
                        hipEvent_t ev;
                        if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
                        b->tev.push_back(ev);
                    }
                }
                t0 = b->tev[b->tevNext];
                t1 = b->tev[b->tevNext + 1];
                b->tevNext += 2;
                if ((e = hipEventRecord(t0, st)) != hipSuccess) return e;
            }
            b->evAfter = timeIt ? t1 : nullptr;   // recorded right after the particle kernel
            b->storeVB = (i == n - 1);
            e = launch_pass(b, true, true, gV, B, i, st, c0, nch);
            b->evAfter = nullptr;
            b->storeVB = false;
            if (e != hipSuccess) return e;
            std::swap(A, B);
            std::swap(fA, fB);
        }
    }
    for (int g = 1; g < groups; ++g) {
        if ((e = hipEventRecord(b->evJoin[g - 1], b->side[g - 1])) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(b->stream, b->evJoin[g - 1], 0)) != hipSuccess) return e;
    }
    *end = cur ^ (n & 1);
    return hipSuccess;
}

}  // namespace

extern "C" {

int icempm_batch_create(icempm_batch **out, int precision) {
    return host_guard(nullptr, [&]() -> int {
        if (!out || (precision != ICEMPM_BATCH_F64 && precision != ICEMPM_BATCH_F32)) return ICEMPM_E_ARG;
        *out = nullptr;
        icempm_batch *b = new icempm_batch();
        b->precision = precision;
        hipError_t e = hipSetDevice(0);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->evFork, hipEventDisableTiming);
        for (int g = 0; g + 1 < icempm_batch::kChains && e == hipSuccess; ++g) {
            e = hipStreamCreateWithFlags(&b->side[g], hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&b->evJoin[g], hipEventDisableTiming);
        }
        if (e == hipSuccess) e = hipEventCreate(&b->e0);
        if (e == hipSuccess) e = hipEventCreate(&b->e1);
        if (e != hipSuccess) {
            icempm_batch_destroy(b);
            return ICEMPM_E_HIP;
        }
        *out = b;
        return ICEMPM_OK;
    });
}

void icempm_batch_destroy(icempm_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipStream_t st : b->side)
        if (st) (void)hipStreamSynchronize(st);
    drop_graph(b);
    dfree(b->pos); dfree(b->rest); dfree(b->crushed); dfree(b->gridA); dfree(b->gridB); dfree(b->vel32);
    dfree(b->blkFlags);
    dfree(b->cases); dfree(b->chunks); dfree(b->nodeBlock);
    dfree(b->err); dfree(b->tTab); dfree(b->hist);
    dfree(b->mlpW); dfree(b->mlpP); dfree(b->hx); dfree(b->ht); dfree(b->hcount); dfree(b->stats);
    dfree(b->mapQ); dfree(b->mapCount);
    dfree(b->dcells);
    if (b->hcells) (void)hipHostFree(b->hcells);
    b->hcells = nullptr;
    dfree(b->orig); dfree(b->sKeys); dfree(b->sVals); dfree(b->sKeys2); dfree(b->sVals2); dfree(b->sTmp);
    dfree(b->tpos); dfree(b->trest); dfree(b->tcrushed); dfree(b->torig);
    dfree(b->kpos); dfree(b->krest); dfree(b->kcrushed); dfree(b->korig);
    dfree(b->grain); dfree(b->tgrain); dfree(b->kgrain); dfree(b->bins);
    for (hipEvent_t ev : b->tev) (void)hipEventDestroy(ev);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->evFork) (void)hipEventDestroy(b->evFork);
    for (hipEvent_t ev : b->evJoin)
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t st : b->side)
        if (st) (void)hipStreamDestroy(st);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *icempm_batch_last_error(const icempm_batch *b) { return b ? b->lastError.c_str() : ""; }

int icempm_batch_add_case(icempm_batch *b, const void *simparams, size_t nbytes, const double *soa,
                          size_t pitch) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !simparams || !soa || nbytes != sizeof(icy::SimParams)) return -ICEMPM_E_ARG;
        if (b->finalized) return -bfail(b, ICEMPM_E_STATE, "add_case: batch already finalized");
        icy::SimParams P;
        std::memcpy(&P, simparams, sizeof(P));
        if (P.nPts <= 0 || (size_t)P.nPts > pitch || P.GridX <= 0 || P.GridY <= 0)
            return -bfail(b, ICEMPM_E_ARG, "add_case: bad sizes");
        if (P.GridX > 32767 || P.GridY > 32767)   // cell indices travel as 16-bit halves in the particle pass
            return -bfail(b, ICEMPM_E_ARG, "add_case: grid wider or taller than 32767 cells");
        // the case's grid slice, its node blocks and every node index the kernels
        // form (X + Y * GridX) are sized from GridTotal: it must be the product
        if ((long)P.GridTotal != (long)P.GridX * (long)P.GridY)
            return -bfail(b, ICEMPM_E_ARG, "add_case: GridTotal is not GridX * GridY");
        // the stencils use cellsize_inv, the weights and stresses cellsize and dt
        // (derived as in ComputeHelperVariables, parameters_sim.cpp:131-133)
        if (!(P.cellsize > 0) || !(P.InitialTimeStep > 0) || !(P.cellsize_inv == 1. / P.cellsize) ||
            !(P.Dp_inv == 4. / (P.cellsize * P.cellsize)))
            return -bfail(b, ICEMPM_E_ARG, "add_case: cellsize, cellsize_inv, Dp_inv and dt must be positive and consistent");
        if (!(P.GrainVariability == P.GrainVariability))
            return -bfail(b, ICEMPM_E_ARG, "add_case: GrainVariability is NaN");
        if (P.n_indenter_subdivisions <= 0 || P.n_indenter_subdivisions > 1 << 16)
            return -bfail(b, ICEMPM_E_ARG, "add_case: n_indenter_subdivisions out of range");
        if (!b->hc.empty() && P.InitialTimeStep != b->hc[0].cp.dt)
            return -bfail(b, ICEMPM_E_ARG, "add_case: all cases of a batch share one time step");
        icempm_batch::HostCase h{};
        CaseParams &c = h.cp;
        c.dt = P.InitialTimeStep; c.vol = P.ParticleVolume; c.mass = P.ParticleMass; c.h = P.cellsize;
        c.h_inv = P.cellsize_inv; c.Dinv = P.Dp_inv; c.mu = P.mu; c.kappa = P.kappa; c.gravity = P.Gravity;
        c.indRsq = P.IndRSq; c.indV = P.IndVelocity; c.ics = P.IceCompressiveStrength;
        c.its = P.IceTensileStrength; c.iss = P.IceShearStrength; c.tanPhi = P.DP_tan_phi;
        c.dpThr = P.DP_threshold_p; c.x0 = P.indenter_x_initial; c.y0 = P.indenter_y;
        {
            // NACC constants of grain 0 with no variability (gpu_implementation5.cu:97-141)
            const double pmax = P.IceCompressiveStrength, pmin = -P.IceTensileStrength, qmax = P.IceShearStrength;
            const double beta = -pmin / pmax;
            const double a = pmax * (1 + beta);
            c.pmax = pmax;
            c.msq = (4 * qmax * qmax * (1 + 2 * beta)) / (a * a);
            c.betaPmax = beta * pmax;
            c.onePlus2Beta = 1. + 2. * beta;
            c.slope = -(2 * std::sqrt(-pmax * pmin) * qmax / (pmax - pmin)) / -3e6;
        }
        c.gx = P.GridX; c.gy = P.GridY; c.gridTotal = P.GridTotal; c.nPts = P.nPts;
        c.gxInv = 1.0 / c.gx;
        c.grainVar = P.GrainVariability;
        c.nbins = P.n_indenter_subdivisions;
        c.binOffset = 0;   // set by finalize
        c.scm = c.dt * c.vol * c.Dinv / c.mass;
        {
            // FP32 storage: 48 - ib fraction bits, ib the bits a cell index needs
            // (finalize refuses grids beyond 65536 nodes a side: ib <= 16)
            int ib = 0;
            while (ib < 16 && (1 << ib) < std::max(c.gx, c.gy)) ++ib;
            const int F = 48 - ib;
            c.cellQ = std::ldexp(1.0, -F);
            c.cellQ16 = std::ldexp(1.0, 16 - F);
            c.cellInvQ = std::ldexp(1.0, F);
            c.posQ = c.h * c.cellQ;
            c.posInvQ = c.cellInvQ / c.h;
        }
        const int n = P.nPts;
        h.pos.resize(2 * (size_t)n);
        h.rest.resize(11 * (size_t)n);
        h.crushed.resize(n);
        h.grain.resize(n);
        h.keys.resize(n);
        const unsigned char *u8 = reinterpret_cast<const unsigned char *>(soa);
        for (int i = 0; i < n; ++i) {
            h.pos[i] = soa[pitch * 4 + i];
            h.pos[n + i] = soa[pitch * 5 + i];
            for (int r = 0; r < 10; ++r) h.rest[(size_t)r * n + i] = soa[pitch * (6 + r) + i];
            h.rest[10 * (size_t)n + i] = soa[pitch * 3 + i];
            h.crushed[i] = u8[i];
            std::memcpy(&h.grain[i], u8 + pitch + 2 * (size_t)i, sizeof(short));   // point.cpp:19-23
            // the base cell's band key, clamped into the grid as on the device
            const double x = h.pos[i], y = h.pos[n + i];
            int key = c.gridTotal - 1;
            if (x == x && y == y) {
                const double bx = std::fmin(std::fmax(x * c.h_inv - 0.5, 0.0), (double)(c.gx - 1));
                const double by = std::fmin(std::fmax(y * c.h_inv - 0.5, 0.0), (double)(c.gy - 1));
                key = band_key((int)bx, (int)by, c.gx, c.gy);
            }
            h.keys[i] = key;
        }
        b->hc.push_back(std::move(h));
        return (int)b->hc.size() - 1;
    });
}

int icempm_batch_finalize(icempm_batch *b) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || b->hc.empty()) return ICEMPM_E_ARG;
        if (b->finalized) return bfail(b, ICEMPM_E_STATE, "finalize: already finalized");
        // point and node offsets, node-block table, keys of the points in slot order
        std::vector<int2> nblk;
        std::vector<int> keys;
        long pofs = 0, gofs = 0, bofs = 0;
        bool anyGrains = false;
        // every case's points in band order (stable): slot o + k holds point
        // order[k] of the case, and orig maps it back
        std::vector<std::vector<int>> order(b->hc.size());
        for (size_t ci = 0; ci < b->hc.size(); ++ci) {
            auto &h = b->hc[ci];
            h.cp.ptOffset = (int)pofs;
            h.cp.gridOffset = (int)gofs;
            h.cp.binOffset = (int)bofs;
            bofs += 2 * (long)h.cp.nbins;
            anyGrains = anyGrains || h.cp.grainVar != 0;
            auto &ord = order[ci];
            ord.resize((size_t)h.cp.nPts);
            for (int i = 0; i < h.cp.nPts; ++i) ord[(size_t)i] = i;
            std::stable_sort(ord.begin(), ord.end(), [&h](int a, int c) { return h.keys[(size_t)a] < h.keys[(size_t)c]; });
            for (int k = 0; k < h.cp.nPts; ++k) keys.push_back(h.keys[(size_t)ord[(size_t)k]]);
            for (int nd = 0; nd < h.cp.gridTotal; nd += 256) nblk.push_back(make_int2((int)ci, nd));
            pofs += h.cp.nPts;
            gofs += h.cp.gridTotal;
        }
        // the particle pass addresses rows through 32-bit byte offsets
        if (pofs >= (1L << 29) || gofs >= (1L << 29) || bofs > INT_MAX / 2)
            return bfail(b, ICEMPM_E_ARG, "finalize: batch too large (2^29 points or grid nodes at most)");
        // a base cell packs its column and row into 16 bits each
        for (const auto &h : b->hc)
            if (h.cp.gx > 65536 || h.cp.gy > 65536)
                return bfail(b, ICEMPM_E_ARG, "finalize: a case's grid is wider or taller than 65536 nodes");
        b->binsTotal = (int)bofs;
        b->nPts = (int)pofs;
        b->nNodes = gofs;
        b->nNodeBlocks = (int)nblk.size();
        // the groups of cases: group g starts at the case boundary nearest to
        // g / chains of all points (each group non-empty)
        {
            const int nc = (int)b->hc.size();
            b->chains = std::max(1, std::min(icempm_batch::kChains, nc));
            b->groupCase[0] = 0;
            long acc = 0;   // points of the cases before case k
            int k = 0;
            for (int g = 1; g < b->chains; ++g) {
                const long target = (long)g * pofs;   // in units of 1 / chains points
                if (k <= b->groupCase[g - 1]) {
                    acc += b->hc[k].cp.nPts;
                    ++k;
                }
                while (k < nc - (b->chains - g) &&
                       std::labs((acc + b->hc[k].cp.nPts) * b->chains - target) <= std::labs(acc * b->chains - target)) {
                    acc += b->hc[k].cp.nPts;
                    ++k;
                }
                b->groupCase[g] = k;
            }
            b->groupCase[b->chains] = nc;
            for (int g = 0; g <= b->chains; ++g) {
                int kb = (int)nblk.size();
                for (size_t q = 0; q < nblk.size(); ++q)
                    if (nblk[q].x >= b->groupCase[g]) {
                        kb = (int)q;
                        break;
                    }
                b->grpBlock[g] = kb;
            }
        }
        b->pitch = (pofs + 63) / 64 * 64;
        b->gpitch = ((size_t)gofs + 63) / 64 * 64;
        {
            // the cells of the band keys, case-local (the key of a cell is
            // band_key of its column and row)
            std::vector<unsigned> cells(keys.size());
            for (size_t ci = 0; ci < b->hc.size(); ++ci) {
                const CaseParams &cp = b->hc[ci].cp;
                for (int k = 0; k < cp.nPts; ++k) {
                    int cx, cy;
                    band_cell(keys[(size_t)cp.ptOffset + k], cp.gx, cp.gy, cx, cy);
                    cells[(size_t)cp.ptOffset + k] = (unsigned)cx | (unsigned)cy << 16;
                }
            }
            chunk_table(b, cells.data(), b->hcstart, b->hccase);
        }
        const size_t vs = b->precision == ICEMPM_BATCH_F32 ? sizeof(float) : sizeof(double);
        BCHK(b, hipMalloc(&b->pos, pos_bytes(b)), "batch finalize");
        BCHK(b, hipMalloc(&b->rest, 11 * b->pitch * vs), "batch finalize");
        BCHK(b, hipMalloc(&b->crushed, b->pitch), "batch finalize");
        if (anyGrains) BCHK(b, hipMalloc(&b->grain, b->pitch * sizeof(short)), "batch finalize");
        BCHK(b, hipMalloc(&b->bins, (size_t)b->binsTotal * sizeof(double)), "batch finalize");
        BCHK(b, hipMalloc(&b->orig, b->pitch * sizeof(int)), "batch finalize");
        BCHK(b, hipMalloc(&b->gridA, 3 * b->gpitch * sizeof(double)), "batch finalize");
        BCHK(b, hipMalloc(&b->gridB, 3 * b->gpitch * sizeof(double)), "batch finalize");
        if (b->precision == ICEMPM_BATCH_F32)
            BCHK(b, hipMalloc(&b->vel32, b->gpitch * sizeof(float2)), "batch finalize");
        BCHK(b, hipMalloc(&b->cases, b->hc.size() * sizeof(CaseParams)), "batch finalize");
        BCHK(b, hipMalloc(&b->nodeBlock, nblk.size() * sizeof(int2)), "batch finalize");
        BCHK(b, hipMalloc(&b->blkFlags, 3 * nblk.size() * sizeof(int)), "batch finalize");
        BCHK(b, hipMalloc(&b->err, sizeof(int)), "batch finalize");
        BCHK(b, hipMalloc(&b->hcount, sizeof(unsigned long long)), "batch finalize");
        BCHK(b, hipMalloc(&b->stats, kMapStats * sizeof(unsigned long long)), "batch finalize");
        // host staging in the device layout
        std::vector<double> hpos(2 * b->pitch, 0.0);
        std::vector<unsigned char> hrest(11 * b->pitch * vs, 0), hcr(b->pitch, 0);
        std::vector<short> hgr(b->pitch, 0);
        std::vector<int> horig(b->pitch, 0);
        std::vector<CaseParams> hcp;
        for (size_t ci = 0; ci < b->hc.size(); ++ci) {
            const auto &h = b->hc[ci];
            const int n = h.cp.nPts, o = h.cp.ptOffset;
            for (int k = 0; k < n; ++k) {
                const int i = order[ci][(size_t)k];
                hpos[o + k] = h.pos[i];
                hpos[b->pitch + o + k] = h.pos[n + i];
                hcr[o + k] = h.crushed[i];
                hgr[o + k] = h.grain[i];
                horig[o + k] = o + i;
                for (int r = 0; r < 11; ++r) {
                    double v = h.rest[(size_t)r * n + i];
                    if (b->precision == ICEMPM_BATCH_F32) {
                        if (r == 2 || r == 5 || r == 10) v -= 1.0;    // Fe00, Fe11, Jp_inv stored less 1
                        float f = (float)v;
                        std::memcpy(&hrest[rest_elem(b, r, (size_t)(o + k)) * vs], &f, vs);
                    } else {
                        std::memcpy(&hrest[rest_elem(b, r, (size_t)(o + k)) * vs], &v, vs);
                    }
                }
            }
            hcp.push_back(h.cp);
        }
        if (b->precision == ICEMPM_BATCH_F32) {
            // one record per slot: 48-bit fixed point (high words of x and y,
            // both low halves) and the crushed flag
            std::vector<unsigned int> q(4 * b->pitch, 0u);
            for (size_t ci = 0; ci < hcp.size(); ++ci) {
                const CaseParams &cp = hcp[ci];
                for (int k = 0; k < cp.nPts; ++k) {
                    const size_t i = (size_t)cp.ptOffset + k;
                    unsigned int hx, lx, hy, ly;
                    fix48_put(hpos[i], cp, hx, lx);
                    fix48_put(hpos[b->pitch + i], cp, hy, ly);
                    q[4 * i] = hx;
                    q[4 * i + 1] = hy;
                    q[4 * i + 2] = lx | (ly << 16);
                    q[4 * i + 3] = hcr[i] | ((unsigned int)(unsigned short)hgr[i] << 16);
                }
            }
            BCHK(b, hipMemcpy(b->pos, q.data(), q.size() * sizeof(unsigned int), hipMemcpyHostToDevice), "batch upload");
        } else {
            BCHK(b, hipMemcpy(b->pos, hpos.data(), hpos.size() * sizeof(double), hipMemcpyHostToDevice), "batch upload");
        }
        BCHK(b, hipMemcpy(b->rest, hrest.data(), hrest.size(), hipMemcpyHostToDevice), "batch upload");
        BCHK(b, hipMemcpy(b->crushed, hcr.data(), hcr.size(), hipMemcpyHostToDevice), "batch upload");
        if (b->grain)
            BCHK(b, hipMemcpy(b->grain, hgr.data(), hgr.size() * sizeof(short), hipMemcpyHostToDevice), "batch upload");
        {
            const std::vector<double> zb((size_t)b->binsTotal, 0.0);
            BCHK(b, hipMemcpy(b->bins, zb.data(), zb.size() * sizeof(double), hipMemcpyHostToDevice), "batch upload");
        }
        BCHK(b, hipMemcpy(b->orig, horig.data(), horig.size() * sizeof(int), hipMemcpyHostToDevice), "batch upload");
        BCHK(b, hipMemcpy(b->cases, hcp.data(), hcp.size() * sizeof(CaseParams), hipMemcpyHostToDevice), "batch upload");
        BCHK(b, hipMemcpy(b->nodeBlock, nblk.data(), nblk.size() * sizeof(int2), hipMemcpyHostToDevice), "batch upload");
        int zero = 0;
        BCHK(b, hipMemcpy(b->err, &zero, sizeof(int), hipMemcpyHostToDevice), "batch upload");
        const unsigned long long z7[kMapStats] = {};
        BCHK(b, hipMemcpy(b->hcount, z7, sizeof(unsigned long long), hipMemcpyHostToDevice), "batch upload");
        BCHK(b, hipMemcpy(b->stats, z7, sizeof(z7), hipMemcpyHostToDevice), "batch upload");
        int rc = upload_chunks(b);
        if (rc) return rc;
        b->finalized = true;
        return ICEMPM_OK;
    });
}

}  // extern "C"

namespace {

// n substeps with the given time table and force-history rows, from the
// P2G the previous run left (b->pending) or from a fresh P2G, as one graph
// launch (captured when the start, the step count or the pointers change)
// or as stream launches; returns the device time in ms.
// The executable graph of n substeps from the given start (-1 fresh, 0 or 1:
// the P2G in grid A or B) with these time-table and history rows, captured,
// instantiated and put on the device unless its slot already holds it.
int ensure_graph(icempm_batch *b, int n, const double *tTab, double *hist, int start,
                 const icempm_batch::GraphSlot **out) {
    icempm_batch::GraphSlot *row = b->graphs[start + 1];
    for (int w = 0; w < 2; ++w)
        if (row[w].exec && row[w].steps == n && row[w].T == tTab && row[w].H == hist) {
            row[w].used = ++b->graphTick;
            if (out) *out = &row[w];
            return ICEMPM_OK;
        }
    icempm_batch::GraphSlot &g = row[row[0].used <= row[1].used ? 0 : 1];
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = icempm_batch::GraphSlot{};
    hipGraph_t graph = nullptr;
    int end = -1;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "batch capture");
    hipError_t e = enqueue_loop(b, n, tTab, hist, start, &end);
    hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return bfail(b, ICEMPM_E_HIP, std::string("batch capture: ") + hipGetErrorString(e));
    }
    BCHK(b, e2, "batch capture");
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    BCHK(b, e, "batch graph instantiate");
    g.steps = n;
    g.T = tTab;
    g.H = hist;
    g.end = end;
    g.used = ++b->graphTick;
    if (out) *out = &g;
    // put the executable graph on the device now, so that its first launch
    // costs what a replay costs
    BCHK(b, hipGraphUpload(g.exec, b->stream), "batch graph upload");
    BCHK(b, hipStreamSynchronize(b->stream), "batch graph upload");
    return ICEMPM_OK;
}

// the time-table and history buffers for runs of up to n substeps
int ensure_capacity(icempm_batch *b, int n) {
    if (n <= b->cap) return ICEMPM_OK;
    drop_graph(b);
    dfree(b->tTab);
    dfree(b->hist);
    b->cap = 0;
    BCHK(b, hipMalloc(&b->tTab, (size_t)n * sizeof(double)), "batch run");
    BCHK(b, hipMalloc(&b->hist, 2 * b->hc.size() * (size_t)n * sizeof(double)), "batch run");
    b->cap = n;
    return ICEMPM_OK;
}

int run_segment(icempm_batch *b, int n, int use_graph, const double *tTab, double *hist, float *ms) {
    b->tevNext = 0;
    const int start = b->pending;
    b->pending = -1;   // until the loop below is enqueued whole
    int end = -1;
    if (use_graph) {
        const icempm_batch::GraphSlot *g = nullptr;
        const int rc = ensure_graph(b, n, tTab, hist, start, &g);
        if (rc) return rc;
        end = g->end;
        // timed from here: capture and instantiation are host work before it
        BCHK(b, hipEventRecord(b->e0, b->stream), "batch run");
        BCHK(b, hipGraphLaunch(g->exec, b->stream), "batch graph launch");
    } else {
        BCHK(b, hipEventRecord(b->e0, b->stream), "batch run");
        BCHK(b, enqueue_loop(b, n, tTab, hist, start, &end), "batch run");
    }
    BCHK(b, hipEventRecord(b->e1, b->stream), "batch run");
    BCHK(b, hipEventSynchronize(b->e1), "batch run");
    b->pending = end;
    BCHK(b, hipEventElapsedTime(ms, b->e0, b->e1), "batch run");
    for (size_t q = 0; q + 1 < b->tevNext; q += 2) {
        float kms = 0;
        BCHK(b, hipEventElapsedTime(&kms, b->tev[q], b->tev[q + 1]), "batch timing");
        b->timedMs += kms;
        b->timedLaunches += 1;
    }
    b->tevNext = 0;
    return ICEMPM_OK;
}

}  // namespace

extern "C" {

int icempm_batch_run(icempm_batch *b, int n, int use_graph) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || n <= 0) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "run: batch not finalized");
        if (use_graph && b->timingStride > 0)
            return bfail(b, ICEMPM_E_STATE, "run: kernel timing needs stream launches (use_graph = 0)");
        const size_t nc = b->hc.size();
        int rc0 = ensure_capacity(b, n);
        if (rc0) return rc0;
        // the time of each substep by repeated addition (model.cpp:23)
        std::vector<double> tt((size_t)n);
        double t = b->t;
        for (int i = 0; i < n; ++i) {
            t += b->hc[0].cp.dt;
            tt[i] = t;
        }
        BCHK(b, hipMemcpyAsync(b->tTab, tt.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream), "batch run");
        // the force bins sum over the run (Model::Step sums them over a cycle, model.cpp:18)
        BCHK(b, launch_fill_f64(b->bins, (size_t)b->binsTotal, 0.0, b->stream), "batch run");
        BCHK(b, hipStreamSynchronize(b->stream), "batch run");
        // segments of resortEvery substeps, the points re-sorted between them
        const int seg = b->resortEvery > 0 ? b->resortEvery : n;
        double total = 0;
        for (int s0 = 0; s0 < n; s0 += seg) {
            if (s0 > 0) {
                const auto h0 = std::chrono::steady_clock::now();
                int rc = resort(b);
                if (rc) return rc;
                total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
            }
            const int len = std::min(seg, n - s0);
            float ms = 0;
            const auto h0 = std::chrono::steady_clock::now();
            int rc = run_segment(b, len, use_graph, b->tTab + s0, b->hist + 2 * nc * (size_t)s0, &ms);
            if (rc) return rc;
            // with re-sorting, each segment is captured anew: count the host time
            total += b->resortEvery > 0 ? std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count()
                                        : (double)ms;
        }
        b->lastMs = (float)total;
        b->t = t;
        b->step += n;
        b->lastRun = n;
        return ICEMPM_OK;
    });
}

int icempm_batch_prepare(icempm_batch *b, int n) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || n <= 0) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "prepare: batch not finalized");
        BCHK(b, hipStreamSynchronize(b->stream), "prepare");
        int rc = ensure_capacity(b, n);
        if (rc) return rc;
        // the first segment of such a run, from each of the three starts
        const int len = b->resortEvery > 0 ? std::min(b->resortEvery, n) : n;
        for (int start = -1; start <= 1; ++start)
            if ((rc = ensure_graph(b, len, b->tTab, b->hist, start, nullptr))) return rc;
        return ICEMPM_OK;
    });
}

int icempm_batch_last_ms(const icempm_batch *b, float *ms) {
    if (!b || !ms) return ICEMPM_E_ARG;
    *ms = b->lastMs;
    return ICEMPM_OK;
}

int icempm_batch_num_points(const icempm_batch *b) { return b ? b->nPts : 0; }

long icempm_batch_num_nodes(const icempm_batch *b) { return b ? b->nNodes : 0; }

int icempm_batch_kernel_timing(icempm_batch *b, int stride) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || stride < 0) return ICEMPM_E_ARG;
        b->timingStride = stride;
        b->timedMs = 0;
        b->timedLaunches = 0;
        return ICEMPM_OK;
    });
}

int icempm_batch_kernel_timing_read(icempm_batch *b, double *total_ms, int *launches) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !total_ms || !launches) return ICEMPM_E_ARG;
        *total_ms = b->timedMs;
        *launches = b->timedLaunches;
        return ICEMPM_OK;
    });
}

int icempm_batch_force_history(icempm_batch *b, int c, double *out, int n) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !out || c < 0 || (size_t)c >= b->hc.size() || n < 0 || n > b->lastRun) return ICEMPM_E_ARG;
        const size_t nc = b->hc.size();
        std::vector<double> all(2 * nc * (size_t)n);
        if (n == 0) return ICEMPM_OK;
        BCHK(b, hipMemcpy(all.data(), b->hist, all.size() * sizeof(double), hipMemcpyDeviceToHost), "force history");
        for (int i = 0; i < n; ++i) {
            out[2 * i] = all[(size_t)i * 2 * nc + 2 * c];
            out[2 * i + 1] = all[(size_t)i * 2 * nc + 2 * c + 1];
        }
        return ICEMPM_OK;
    });
}

int icempm_batch_force_bins(icempm_batch *b, int c, double *out, int nbins) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !out || c < 0 || (size_t)c >= b->hc.size()) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "force_bins: batch not finalized");
        const CaseParams &cp = b->hc[c].cp;
        if (nbins != cp.nbins) return bfail(b, ICEMPM_E_ARG, "force_bins: nbins is not the case's n_indenter_subdivisions");
        BCHK(b, hipStreamSynchronize(b->stream), "force_bins");
        BCHK(b, hipMemcpy(out, b->bins + cp.binOffset, 2 * (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost),
             "force_bins");
        return ICEMPM_OK;
    });
}

int icempm_batch_read_points(icempm_batch *b, int c, double *soa, size_t pitch) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !soa || c < 0 || (size_t)c >= b->hc.size()) return ICEMPM_E_ARG;
        const CaseParams &cp = b->hc[c].cp;
        const int n = cp.nPts, o = cp.ptOffset;
        if ((size_t)n > pitch) return bfail(b, ICEMPM_E_ARG, "read_points: pitch too small");
        const size_t vs = b->precision == ICEMPM_BATCH_F32 ? sizeof(float) : sizeof(double);
        std::vector<double> hpos(2 * (size_t)n);
        std::vector<unsigned char> hrest(11 * (size_t)n * vs), hcr(n);
        std::vector<short> hgr((size_t)n, 0);
        if (b->precision == ICEMPM_BATCH_F32) {
            std::vector<unsigned int> q(4 * (size_t)n);
            const unsigned int *P = static_cast<const unsigned int *>(b->pos);
            BCHK(b, hipMemcpy(q.data(), P + 4 * (size_t)o, 4 * (size_t)n * sizeof(unsigned int), hipMemcpyDeviceToHost),
                 "read_points");
            for (int i = 0; i < n; ++i) {
                const unsigned int lo = q[4 * (size_t)i + 2];
                hpos[i] = fix48_get(q[4 * (size_t)i], lo & 0xFFFFu, cp);
                hpos[n + i] = fix48_get(q[4 * (size_t)i + 1], lo >> 16, cp);
                hcr[i] = (unsigned char)(q[4 * (size_t)i + 3] & 0xFFu);
                hgr[i] = (short)(q[4 * (size_t)i + 3] >> 16);
            }
        } else {
            const double *P = static_cast<const double *>(b->pos);
            BCHK(b, hipMemcpy(hpos.data(), P + o, n * sizeof(double), hipMemcpyDeviceToHost), "read_points");
            BCHK(b, hipMemcpy(hpos.data() + n, P + b->pitch + o, n * sizeof(double), hipMemcpyDeviceToHost),
                 "read_points");
        }
        // rows of the case's slots [o, o + n) into hrest[r][slot]; FP32 storage's
        // Fe records come down as one block and are split
        for (int r = 0; r < 11; ++r) {
            if (b->precision == ICEMPM_BATCH_F32 && r >= 2 && r <= 5) continue;
            BCHK(b, hipMemcpy(hrest.data() + (size_t)r * n * vs, (unsigned char *)b->rest + rest_elem(b, r, o) * vs,
                              n * vs, hipMemcpyDeviceToHost), "read_points");
        }
        if (b->precision == ICEMPM_BATCH_F32) {
            std::vector<float> fe(4 * (size_t)n);
            BCHK(b, hipMemcpy(fe.data(), (unsigned char *)b->rest + rest_elem(b, 2, o) * vs, 4 * (size_t)n * vs,
                              hipMemcpyDeviceToHost), "read_points");
            for (int slot = 0; slot < n; ++slot)
                for (int r = 2; r <= 5; ++r)
                    std::memcpy(&hrest[((size_t)r * n + slot) * vs], &fe[4 * (size_t)slot + (r - 2)], vs);
        }
        if (b->precision != ICEMPM_BATCH_F32)   // FP32: from the position records above
            BCHK(b, hipMemcpy(hcr.data(), b->crushed + o, n, hipMemcpyDeviceToHost), "read_points");
        if (b->grain && b->precision != ICEMPM_BATCH_F32)
            BCHK(b, hipMemcpy(hgr.data(), b->grain + o, n * sizeof(short), hipMemcpyDeviceToHost), "read_points");
        std::vector<int> horig((size_t)n);
        BCHK(b, hipMemcpy(horig.data(), b->orig + o, n * sizeof(int), hipMemcpyDeviceToHost), "read_points");
        std::memset(soa, 0, 16 * pitch * sizeof(double));
        unsigned char *u8 = reinterpret_cast<unsigned char *>(soa);
        for (int slot = 0; slot < n; ++slot) {
            // the point in this slot, at its index within the case
            const int i = horig[slot] - o;
            if (i < 0 || i >= n) return bfail(b, ICEMPM_E_STATE, "read_points: slot map corrupt");
            u8[i] = hcr[slot];
            std::memcpy(u8 + pitch + 2 * (size_t)i, &hgr[slot], sizeof(short));   // point.cpp:19-23
            soa[pitch * 4 + i] = hpos[slot];
            soa[pitch * 5 + i] = hpos[n + slot];
            for (int r = 0; r < 11; ++r) {
                double v;
                if (vs == sizeof(float)) {
                    float f;
                    std::memcpy(&f, &hrest[((size_t)r * n + slot) * vs], vs);
                    v = f;
                    if (r == 2 || r == 5 || r == 10) v += 1.0;
                } else {
                    std::memcpy(&v, &hrest[((size_t)r * n + slot) * vs], vs);
                }
                if (r < 10)
                    soa[pitch * (6 + r) + i] = v;
                else
                    soa[pitch * 3 + i] = v;
            }
        }
        return ICEMPM_OK;
    });
}

int icempm_batch_error_flag(icempm_batch *b, int *flag) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !flag) return ICEMPM_E_ARG;
        BCHK(b, hipMemcpy(flag, b->err, sizeof(int), hipMemcpyDeviceToHost), "error flag");
        return ICEMPM_OK;
    });
}

int icempm_batch_set_return_map(icempm_batch *b, int mode, const float *params, const float *norm,
                                double tol, int harvest_capacity) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "set_return_map: batch not finalized");
        if (mode != kMapExact && mode != kMapHarvest && mode != kMapLearned)
            return bfail(b, ICEMPM_E_ARG, "set_return_map: unknown mode");
        if (mode == kMapHarvest && harvest_capacity <= 0)
            return bfail(b, ICEMPM_E_ARG, "set_return_map: harvest needs a positive capacity");
        if (mode == kMapLearned && (!params || !norm))
            return bfail(b, ICEMPM_E_ARG, "set_return_map: the learned map needs parameters and a normalisation");
        if (mode == kMapLearned && !(tol == tol))
            return bfail(b, ICEMPM_E_ARG, "set_return_map: tolerance is NaN");
        BCHK(b, hipStreamSynchronize(b->stream), "set_return_map");
        // a captured loop holds the previous kernel arguments
        drop_graph(b);
        if (mode == kMapHarvest) {
            if (harvest_capacity != b->hcap) {
                dfree(b->hx);
                dfree(b->ht);
                b->hcap = 0;
                BCHK(b, hipMalloc(&b->hx, (size_t)harvest_capacity * 5 * sizeof(float)), "set_return_map");
                BCHK(b, hipMalloc(&b->ht, (size_t)harvest_capacity * 3 * sizeof(float)), "set_return_map");
                b->hcap = harvest_capacity;
            }
            const unsigned long long zero = 0;
            BCHK(b, hipMemcpy(b->hcount, &zero, sizeof(zero), hipMemcpyHostToDevice), "set_return_map");
        }
        if (mode == kMapLearned) {
            const int np = mlpdev::kParams;
            std::vector<unsigned short> wb(np);
            for (int i = 0; i < np; ++i) wb[i] = mlpdev::to_bf16(params[i]);
            if (!b->mlpW) BCHK(b, hipMalloc(&b->mlpW, np * sizeof(unsigned short)), "set_return_map");
            if (!b->mlpP) BCHK(b, hipMalloc(&b->mlpP, np * sizeof(float)), "set_return_map");
            BCHK(b, hipMemcpy(b->mlpW, wb.data(), np * sizeof(unsigned short), hipMemcpyHostToDevice), "set_return_map");
            BCHK(b, hipMemcpy(b->mlpP, params, np * sizeof(float), hipMemcpyHostToDevice), "set_return_map");
            std::memcpy(b->norm.in_mean, norm, 5 * sizeof(float));
            std::memcpy(b->norm.in_scale, norm + 5, 5 * sizeof(float));
            std::memcpy(b->norm.out_mean, norm + 10, 3 * sizeof(float));
            std::memcpy(b->norm.out_scale, norm + 13, 3 * sizeof(float));
            for (int q = 0; q < 5; ++q)
                if (!(b->norm.in_scale[q] > 0)) return bfail(b, ICEMPM_E_ARG, "set_return_map: input scale must be > 0");
            b->tol = tol;
            if (!b->mapQ) BCHK(b, hipMalloc(&b->mapQ, (size_t)b->nPts * sizeof(MapEntry)), "set_return_map");
            if (!b->mapCount) BCHK(b, hipMalloc(&b->mapCount, 2 * sizeof(unsigned int)), "set_return_map");
        }
        b->rmap = mode;
        return ICEMPM_OK;
    });
}

int icempm_batch_read_harvest(icempm_batch *b, float *X, float *T, int cap, unsigned long long *total) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !total || cap < 0 || (cap > 0 && (!X || !T))) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "read_harvest: batch not finalized");
        BCHK(b, hipStreamSynchronize(b->stream), "read_harvest");
        unsigned long long n = 0;
        BCHK(b, hipMemcpy(&n, b->hcount, sizeof(n), hipMemcpyDeviceToHost), "read_harvest");
        *total = n;
        const int m = (int)std::min<unsigned long long>(std::min<unsigned long long>(n, (unsigned long long)b->hcap),
                                                        (unsigned long long)cap);
        if (m > 0) {
            BCHK(b, hipMemcpy(X, b->hx, (size_t)m * 5 * sizeof(float), hipMemcpyDeviceToHost), "read_harvest");
            BCHK(b, hipMemcpy(T, b->ht, (size_t)m * 3 * sizeof(float), hipMemcpyDeviceToHost), "read_harvest");
        }
        return ICEMPM_OK;
    });
}

int icempm_batch_set_tuning(icempm_batch *b, int knob, int value) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b) return ICEMPM_E_ARG;
        switch (knob) {
        case ICEMPM_BATCH_TUNE_CHUNK:
            if (value != 64 && value != 128 && value != 256)
                return bfail(b, ICEMPM_E_ARG, "set_tuning: points per chunk is 64, 128 or 256");
            if (value != b->chunkPoints) {
                b->chunkPoints = value;
                if (b->finalized) {
                    BCHK(b, hipStreamSynchronize(b->stream), "set_tuning");
                    return rebuild_chunks(b);
                }
            }
            return ICEMPM_OK;
        case ICEMPM_BATCH_TUNE_RESORT:
            if (value < 0) return bfail(b, ICEMPM_E_ARG, "set_tuning: the re-sort cadence is >= 0");
            b->resortEvery = value;
            return ICEMPM_OK;
        case ICEMPM_BATCH_TUNE_MAP_MIN_ACCEPT:
            if (value < 0 || value > 100) return bfail(b, ICEMPM_E_ARG, "set_tuning: acceptance threshold is 0..100 percent");
            b->mapMinAccept = value / 100.0;
            return ICEMPM_OK;
        default:
            return bfail(b, ICEMPM_E_ARG, "set_tuning: unknown knob");
        }
    });
}

int icempm_batch_get_tuning(const icempm_batch *b, int knob) {
    if (!b) return -1;
    switch (knob) {
    case ICEMPM_BATCH_TUNE_CHUNK: return b->chunkPoints;
    case ICEMPM_BATCH_TUNE_RESORT: return b->resortEvery;
    case ICEMPM_BATCH_TUNE_MAP_MIN_ACCEPT: return (int)std::lround(b->mapMinAccept * 100.0);
    default: return -1;
    }
}

int icempm_batch_resort(icempm_batch *b) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "resort: batch not finalized");
        BCHK(b, hipStreamSynchronize(b->stream), "resort");
        return resort(b);
    });
}

int icempm_batch_checkpoint(icempm_batch *b) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "checkpoint: batch not finalized");
        if (!b->kpos) {
            BCHK(b, hipMalloc(&b->kpos, pos_bytes(b)), "checkpoint");
            BCHK(b, hipMalloc(&b->krest, rest_bytes(b)), "checkpoint");
            BCHK(b, hipMalloc(&b->kcrushed, b->pitch), "checkpoint");
            BCHK(b, hipMalloc(&b->korig, b->pitch * sizeof(int)), "checkpoint");
            if (b->grain) BCHK(b, hipMalloc(&b->kgrain, b->pitch * sizeof(short)), "checkpoint");
        }
        BCHK(b, hipMemcpyAsync(b->kpos, b->pos, pos_bytes(b), hipMemcpyDeviceToDevice, b->stream), "checkpoint");
        BCHK(b, hipMemcpyAsync(b->krest, b->rest, rest_bytes(b), hipMemcpyDeviceToDevice, b->stream), "checkpoint");
        BCHK(b, hipMemcpyAsync(b->kcrushed, b->crushed, b->pitch, hipMemcpyDeviceToDevice, b->stream), "checkpoint");
        BCHK(b, hipMemcpyAsync(b->korig, b->orig, b->pitch * sizeof(int), hipMemcpyDeviceToDevice, b->stream), "checkpoint");
        if (b->grain)   // re-sorts permute the grain row with the points
            BCHK(b, hipMemcpyAsync(b->kgrain, b->grain, b->pitch * sizeof(short), hipMemcpyDeviceToDevice, b->stream),
                 "checkpoint");
        BCHK(b, hipStreamSynchronize(b->stream), "checkpoint");
        b->kt = b->t;
        b->kstep = b->step;
        b->kcstart = b->hcstart;
        b->kccase = b->hccase;
        b->kChunkPoints = b->chunkPoints;
        b->haveCheckpoint = true;
        return ICEMPM_OK;
    });
}

int icempm_batch_restore(icempm_batch *b) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b) return ICEMPM_E_ARG;
        if (!b->haveCheckpoint) return bfail(b, ICEMPM_E_STATE, "restore: no checkpoint");
        BCHK(b, hipStreamSynchronize(b->stream), "restore");
        BCHK(b, hipMemcpyAsync(b->pos, b->kpos, pos_bytes(b), hipMemcpyDeviceToDevice, b->stream), "restore");
        BCHK(b, hipMemcpyAsync(b->rest, b->krest, rest_bytes(b), hipMemcpyDeviceToDevice, b->stream), "restore");
        BCHK(b, hipMemcpyAsync(b->crushed, b->kcrushed, b->pitch, hipMemcpyDeviceToDevice, b->stream), "restore");
        BCHK(b, hipMemcpyAsync(b->orig, b->korig, b->pitch * sizeof(int), hipMemcpyDeviceToDevice, b->stream), "restore");
        if (b->grain)
            BCHK(b, hipMemcpyAsync(b->grain, b->kgrain, b->pitch * sizeof(short), hipMemcpyDeviceToDevice, b->stream),
                 "restore");
        BCHK(b, hipStreamSynchronize(b->stream), "restore");
        b->pending = -1;   // the grids hold the P2G of the state before the restore
        b->t = b->kt;
        b->step = b->kstep;
        b->hcstart = b->kcstart;
        b->hccase = b->kccase;
        b->chunkPoints = b->kChunkPoints;
        return upload_chunks(b);
    });
}

int icempm_batch_autotune(icempm_batch *b, int probe_steps, double *report, int cap, int *rows) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || probe_steps < 16 || cap < 0 || (cap > 0 && !report) || !rows) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "autotune: batch not finalized");
        if (b->timingStride > 0) return bfail(b, ICEMPM_E_STATE, "autotune: switch kernel timing off first");
        // the probes would append their pairs to the harvest buffer
        if (b->rmap == kMapHarvest) return bfail(b, ICEMPM_E_STATE, "autotune: not in harvest mode");
        int rc = icempm_batch_checkpoint(b);
        if (rc) return rc;
        // what the probes change besides the point state: the map statistics
        // and the error flag (restored below), the force history and lastRun
        // (the probes' own; a caller reads them after its next run)
        unsigned long long stats0[kMapStats] = {};
        int err0 = 0;
        BCHK(b, hipMemcpy(stats0, b->stats, sizeof(stats0), hipMemcpyDeviceToHost), "autotune");
        BCHK(b, hipMemcpy(&err0, b->err, sizeof(int), hipMemcpyDeviceToHost), "autotune");
        const int chunks[3] = {256, 128, 64};
        const int resorts[3] = {0, probe_steps / 2, probe_steps / 4};
        double best = 1e300;
        int bestChunk = b->chunkPoints, bestResort = b->resortEvery, nrow = 0;
        for (int ci = 0; ci < 3; ++ci)
            for (int ri = 0; ri < 3; ++ri) {
                // one clock for every candidate: the host time of the whole
                // probe (re-sorts and graph captures included), measured on a
                // second run so that the no-re-sort candidates replay a graph
                // captured by the first
                b->chunkPoints = chunks[ci];
                b->resortEvery = resorts[ri];
                double per = 0;
                for (int pass = 0; pass < 2; ++pass) {
                    if ((rc = icempm_batch_restore(b))) return rc;
                    b->chunkPoints = chunks[ci];
                    if ((rc = rebuild_chunks(b))) return rc;
                    const auto h0 = std::chrono::steady_clock::now();
                    if ((rc = icempm_batch_run(b, probe_steps, 1))) return rc;
                    per = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count() /
                          probe_steps;
                }
                if (nrow < cap) {
                    report[3 * nrow] = chunks[ci];
                    report[3 * nrow + 1] = resorts[ri];
                    report[3 * nrow + 2] = per;
                }
                ++nrow;
                if (per < best) {
                    best = per;
                    bestChunk = chunks[ci];
                    bestResort = resorts[ri];
                }
            }
        *rows = nrow;
        // With the learned map: a branch whose network proposals the yield gate
        // passed less often than mapMinAccept over the probes is switched off
        // (its projections take the exact map in the particle pass, with no
        // queue and no map kernel)
        if (b->rmap == kMapLearned) {
            unsigned long long st[kMapStats];
            BCHK(b, hipMemcpy(st, b->stats, sizeof(st), hipMemcpyDeviceToHost), "autotune");
            int mask = 0;
            const unsigned long long tA = st[3] - stats0[3], tF = st[4] - stats0[4];
            const unsigned long long sA = st[5] - stats0[5], sF = st[6] - stats0[6];
            if ((b->mapBranches & kMapTension) && tA + tF > 0 && (double)tA >= b->mapMinAccept * (double)(tA + tF))
                mask |= kMapTension;
            if ((b->mapBranches & kMapShear) && sA + sF > 0 && (double)sA >= b->mapMinAccept * (double)(sA + sF))
                mask |= kMapShear;
            b->tuneBranchStats[0] = tA; b->tuneBranchStats[1] = tF;
            b->tuneBranchStats[2] = sA; b->tuneBranchStats[3] = sF;
            if ((rc = icempm_batch_set_map_branches(b, mask))) return rc;
        }
        // back to the checkpointed state, with the fastest setting
        if ((rc = icempm_batch_restore(b))) return rc;
        BCHK(b, hipMemcpy(b->stats, stats0, sizeof(stats0), hipMemcpyHostToDevice), "autotune");
        BCHK(b, hipMemcpy(b->err, &err0, sizeof(int), hipMemcpyHostToDevice), "autotune");
        b->resortEvery = bestResort;
        if (bestChunk == b->chunkPoints) return ICEMPM_OK;   // the restored table is the one
        b->chunkPoints = bestChunk;
        return rebuild_chunks(b);
    });
}

int icempm_batch_map_stats(icempm_batch *b, unsigned long long *out, int reset) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !out) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "map_stats: batch not finalized");
        BCHK(b, hipStreamSynchronize(b->stream), "map_stats");
        unsigned long long all[kMapStats];
        BCHK(b, hipMemcpy(all, b->stats, sizeof(all), hipMemcpyDeviceToHost), "map_stats");
        std::memcpy(out, all, 3 * sizeof(unsigned long long));
        if (reset) {
            const unsigned long long z[kMapStats] = {};
            BCHK(b, hipMemcpy(b->stats, z, sizeof(z), hipMemcpyHostToDevice), "map_stats");
        }
        return ICEMPM_OK;
    });
}

int icempm_batch_map_branch_stats(icempm_batch *b, unsigned long long *out, int reset) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || !out) return ICEMPM_E_ARG;
        if (!b->finalized) return bfail(b, ICEMPM_E_STATE, "map_branch_stats: batch not finalized");
        BCHK(b, hipStreamSynchronize(b->stream), "map_branch_stats");
        unsigned long long all[kMapStats];
        BCHK(b, hipMemcpy(all, b->stats, sizeof(all), hipMemcpyDeviceToHost), "map_branch_stats");
        std::memcpy(out, all + 3, 4 * sizeof(unsigned long long));
        if (reset) {
            const unsigned long long z[kMapStats] = {};
            BCHK(b, hipMemcpy(b->stats, z, sizeof(z), hipMemcpyHostToDevice), "map_branch_stats");
        }
        return ICEMPM_OK;
    });
}

int icempm_batch_set_map_branches(icempm_batch *b, int mask) {
    return host_guard(b ? &b->lastError : nullptr, [&]() -> int {
        if (!b || mask < 0 || mask > (kMapTension | kMapShear)) return ICEMPM_E_ARG;
        if (mask != b->mapBranches) {
            BCHK(b, hipStreamSynchronize(b->stream), "set_map_branches");
            drop_graph(b);   // the captured launches hold the mask (and the map kernel or not)
            b->mapBranches = mask;
        }
        return ICEMPM_OK;
    });
}

int icempm_batch_get_map_branches(const icempm_batch *b) { return b ? b->mapBranches : -1; }

int icempm_batch_autotune_branch_stats(const icempm_batch *b, unsigned long long *out) {
    if (!b || !out) return ICEMPM_E_ARG;
    std::memcpy(out, b->tuneBranchStats, sizeof(b->tuneBranchStats));
    return ICEMPM_OK;
}

}  // extern "C"
