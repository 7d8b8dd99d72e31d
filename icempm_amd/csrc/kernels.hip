// kernels.hip -- grid kernels of the MLS-MPM substep for gfx950 (CDNA4),
// FP64, built without FMA contraction.  (The particle pass is in
// particles.inc, built twice: particles_exact.hip and particles_fast.hip.)
//
//   k_update<false>  grid update (FAST mode) with the indenter force reduced
//       per workgroup in shared memory; also clears the next substep's grid.
// ORDERED path (bit-reproducible):
//   k_p2g_contrib / k_p2g_fill / k_p2g_fold  every node sums its
//       contributions in increasing particle index;
//   k_update<true> + k_bins_fold  the indenter force is summed in increasing
//       node index.
#include <climits>

#include "internal.h"

namespace icempm {

__global__ void k_fill_f64(double *__restrict__ p, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_i32(int *__restrict__ p, size_t n, int v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static unsigned fill_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? (b == 0 ? 1 : b) : 4096);
}

hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(fill_blocks(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, size_t n, int v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_blocks(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// Plain streaming copy: the reference point for the achievable HBM bandwidth
// and for scaling the HBM traffic counters.  Each thread moves four
// independent 16-byte vectors per iteration so that enough loads are in
// flight per wavefront.
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src,
                                              double2 *__restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 a = src[i], b = src[i + stride], c = src[i + 2 * stride],
                      d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(const double *src, double *dst, size_t n_doubles, hipStream_t s) {
    const size_t n2 = n_doubles / 2;
    if (n2 == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(src), reinterpret_cast<double2 *>(dst), n2);
    return hipGetLastError();
}

__global__ void k_hello(int *flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) flag[0] = 1;
}

hipError_t launch_hello(int *flag, hipStream_t s) {
    hipLaunchKernelGGL(k_hello, dim3(1), dim3(64), 0, s, flag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grid update: v2_kernel_update_nodes (gpu_implementation5.cu:215-290).
constexpr int kUpdBlock = 256;
constexpr int kMaxLdsBins = 256;

template <bool ORDERED>
__global__ __launch_bounds__(kUpdBlock) void k_update(KParams k, double *__restrict__ grid,
                                                      double ind_x, double ind_y,
                                                      double *__restrict__ hist2,
                                                      double *__restrict__ zeroGrid,
                                                      int *__restrict__ nodeBin,
                                                      double *__restrict__ nodeF) {
    __shared__ double sb[2 * kMaxLdsBins + 2];
    __shared__ int any;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = idx < k.gridTotal;
    const size_t gp = k.gp;
    const int nb = k.nbins;
    const bool ldsBins = nb <= kMaxLdsBins;
    if constexpr (!ORDERED) {
        if (ldsBins)
            for (int e = threadIdx.x; e < 2 * nb + 2; e += blockDim.x) sb[e] = 0;
        if (threadIdx.x == 0) any = 0;
        __syncthreads();
    }
    int bin = -1;
    double f0 = 0, f1 = 0;
    if (in) {
        if (zeroGrid) {
            zeroGrid[idx] = 0;
            zeroGrid[gp + idx] = 0;
            zeroGrid[2 * gp + idx] = 0;
        }
        const double mass = grid[idx];
        if (mass != 0) {
            const double dt = k.dt, h = k.h, V = k.indV;
            const int gx = k.gridX, gy = k.gridY;
            double v0 = grid[gp + idx], v1 = grid[2 * gp + idx];
            v0 = v0 / mass;
            v1 = v1 / mass;
            v1 = v1 - dt * k.gravity;
            const double vmax = 0.5 * h / dt;
            const double vn = sqrt(v0 * v0 + v1 * v1);
            if (vn > vmax) {
                v0 = v0 / vn * vmax;
                v1 = v1 / vn * vmax;
            }
            const int ix = idx % gx;
            const int iy = idx / gx;
            double n0 = ix * h - ind_x;
            double n1 = iy * h - ind_y;
            if (n0 * n0 + n1 * n1 < k.indRsq) {
                const double r0 = v0 - V;
                const double r1 = v1 - 0.0;
                const double z = n0 * n0 + n1 * n1;
                if (z > 0) {
                    const double nz = sqrt(z);
                    n0 = n0 / nz;
                    n1 = n1 / nz;
                }
                const double vnr = r0 * n0 + r1 * n1;
                if (vnr < 0) {
                    const double t0 = r0 - n0 * vnr;
                    const double t1 = r1 - n1 * vnr;
                    const double p0 = v0, p1 = v1;
                    v0 = V + t0;
                    v1 = 0.0 + t1;
                    f0 = (p0 - v0) * mass / dt;
                    f1 = (p1 - v1) * mass / dt;
                    double angle = atan2(n0, n1);
                    angle += kPi;
                    angle *= nb / (2 * kPi);
                    int index = (int)angle;
                    index = max(index, 0);
                    index = min(index, nb - 1);
                    bin = index;
                }
            }
            if (iy <= 2) {
                v0 = 0;
                v1 = 0;
            } else if (iy >= gy - 3 && v1 > 0) {
                v1 = 0;
            }
            if (ix <= 2 && v0 < 0)
                v0 = 0;
            else if (ix >= gx - 3 && v0 > 0)
                v0 = 0;
            grid[gp + idx] = v0;
            grid[2 * gp + idx] = v1;
        }
    }
    if constexpr (ORDERED) {
        if (in) {
            nodeBin[idx] = bin;
            nodeF[2 * (size_t)idx] = f0;
            nodeF[2 * (size_t)idx + 1] = f1;
        }
    } else {
        if (bin >= 0) {
            if (ldsBins) {
                atomicAdd(&sb[2 * bin], f0);
                atomicAdd(&sb[2 * bin + 1], f1);
                atomicAdd(&sb[2 * nb], f0);
                atomicAdd(&sb[2 * nb + 1], f1);
                any = 1;
            } else {
                unsafeAtomicAdd(&k.bins[2 * bin], f0);
                unsafeAtomicAdd(&k.bins[2 * bin + 1], f1);
                if (hist2) {
                    unsafeAtomicAdd(&hist2[0], f0);
                    unsafeAtomicAdd(&hist2[1], f1);
                }
            }
        }
        __syncthreads();
        if (ldsBins && any) {
            for (int e = threadIdx.x; e < 2 * nb + 2; e += blockDim.x) {
                const double v = sb[e];
                if (v != 0) {
                    if (e < 2 * nb)
                        unsafeAtomicAdd(&k.bins[e], v);
                    else if (hist2)
                        unsafeAtomicAdd(&hist2[e - 2 * nb], v);
                }
            }
        }
    }
}

hipError_t launch_update(const KParams &k, double *grid, double ix, double iy, double *hist2,
                         double *zeroGrid, bool ordered, int *nodeBin, double *nodeF,
                         hipStream_t s) {
    if (k.gridTotal <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((k.gridTotal + kUpdBlock - 1) / kUpdBlock);
    if (ordered)
        hipLaunchKernelGGL(k_update<true>, dim3(nb), dim3(kUpdBlock), 0, s, k, grid, ix, iy,
                           hist2, zeroGrid, nodeBin, nodeF);
    else
        hipLaunchKernelGGL(k_update<false>, dim3(nb), dim3(kUpdBlock), 0, s, k, grid, ix, iy,
                           hist2, zeroGrid, nodeBin, nodeF);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ORDERED P2G
__global__ __launch_bounds__(256) void k_p2g_contrib(KParams k, double *__restrict__ contrib,
                                                    int *__restrict__ cnode,
                                                    int *__restrict__ count) {
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= k.nPts) return;
    const double *P = k.pts;
    const size_t pp = k.pp;
    const int gx = k.gridX, gy = k.gridY;
    const double px = P[pp * kRowPosX + pt], py = P[pp * (kRowPosX + 1) + pt];
    const double vx = P[pp * kRowVelX + pt], vy = P[pp * (kRowVelX + 1) + pt];
    const double F00 = P[pp * (kRowFe00 + 0) + pt], F01 = P[pp * (kRowFe00 + 1) + pt];
    const double F10 = P[pp * (kRowFe00 + 2) + pt], F11 = P[pp * (kRowFe00 + 3) + pt];
    const double B00 = P[pp * (kRowBp00 + 0) + pt], B01 = P[pp * (kRowBp00 + 1) + pt];
    const double B10 = P[pp * (kRowBp00 + 2) + pt], B11 = P[pp * (kRowBp00 + 3) + pt];
    int *cn = cnode + 9 * (size_t)pt;
    int bx, by;
    double fx0, fx1, wx0, wx1, wx2, wy0, wy1, wy2;
    if (!stencil(px, k.h_inv, bx, fx0, wx0, wx1, wx2) ||
        !stencil(py, k.h_inv, by, fx1, wy0, wy1, wy2)) {
        flag_error(k);
        for (int e = 0; e < 9; ++e) cn[e] = -1;
        return;
    }
    double t00, t01, t11;
    kirchhoff(F00, F01, F10, F11, k.mu, k.kappa, t00, t01, t11);
    const double m = k.mass, h = k.h;
    const double sc = k.dt * k.vol * k.Dinv;
    const double A00 = m * B00 - sc * t00;
    const double A01 = m * B01 - sc * t01;
    const double A10 = m * B10 - sc * t01;
    const double A11 = m * B11 - sc * t11;
    const double mv0 = vx * m, mv1 = vy * m;
    double *co = contrib + 27 * (size_t)pt;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int e = i * 3 + j;
            const double Wip = sel3(i, wx0, wx1, wx2) * sel3(j, wy0, wy1, wy2);
            const double d0 = (i - fx0) * h;
            const double d1 = (j - fx1) * h;
            const double incV0 = Wip * (mv0 + (A00 * d0 + A01 * d1));
            const double incV1 = Wip * (mv1 + (A10 * d0 + A11 * d1));
            const double incM = Wip * m;
            const int X = bx + i, Y = by + j;
            if (X < 0 || Y < 0 || X >= gx || Y >= gy) {
                flag_error(k);
                cn[e] = -1;
                continue;
            }
            const int node = X + Y * gx;
            cn[e] = node;
            co[3 * e + 0] = incM;
            co[3 * e + 1] = incV0;
            co[3 * e + 2] = incV1;
            atomicAdd(&count[node], 1);
        }
    }
}

__global__ void k_p2g_fill(const int *__restrict__ cnode, const int *__restrict__ start,
                           int *__restrict__ cursor, int *__restrict__ list, size_t n9) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n9) return;
    const int node = cnode[e];
    if (node < 0) return;
    const int slot = start[node] + atomicAdd(&cursor[node], 1);
    list[slot] = (int)e;
}

__global__ void k_p2g_fold(double *__restrict__ grid, size_t gp, const int *__restrict__ start,
                           int *__restrict__ list, const double *__restrict__ contrib,
                           int gridTotal) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= gridTotal) return;
    const int b = start[node], e = start[node + 1];
    // insertion sort: the entries are (particle*9 + stencil slot), so the
    // ascending order is the particle order
    for (int i = b + 1; i < e; ++i) {
        const int v = list[i];
        int j = i - 1;
        while (j >= b && list[j] > v) {
            list[j + 1] = list[j];
            --j;
        }
        list[j + 1] = v;
    }
    double m = 0.0, x = 0.0, y = 0.0;
    for (int i = b; i < e; ++i) {
        const double *c = contrib + 3 * (size_t)list[i];
        m += c[0];
        x += c[1];
        y += c[2];
    }
    grid[node] = m;
    grid[gp + node] = x;
    grid[2 * gp + node] = y;
}

hipError_t ordered_p2g(const KParams &k, double *grid, OrderedScratch &o, hipStream_t s) {
    hipError_t e;
    const size_t G = (size_t)k.gridTotal, N = (size_t)k.nPts;
    if ((e = launch_fill_i32(o.count, G + 1, 0, s)) != hipSuccess) return e;
    if ((e = launch_fill_i32(o.cursor, G, 0, s)) != hipSuccess) return e;
    if (N > 0) {
        hipLaunchKernelGGL(k_p2g_contrib, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, k,
                           o.contrib, o.cnode, o.count);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = exclusive_scan(o.count, o.start, G + 1, o.scanTmp, s)) != hipSuccess) return e;
    if (N > 0) {
        hipLaunchKernelGGL(k_p2g_fill, dim3((unsigned)((9 * N + 255) / 256)), dim3(256), 0, s,
                           o.cnode, o.start, o.cursor, o.list, 9 * N);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_p2g_fold, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, grid,
                       k.gp, o.start, o.list, o.contrib, k.gridTotal);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ORDERED indenter force: contact nodes compacted in node order, then each
// bin (and the substep total) folded sequentially.
__global__ void k_contact_flags(const int *__restrict__ nodeBin, int *__restrict__ flags, int G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) flags[i] = nodeBin[i] >= 0 ? 1 : 0;
    if (i == G) flags[i] = 0;
}

__global__ void k_contact_compact(const int *__restrict__ flags, const int *__restrict__ cpos,
                                  int *__restrict__ contacts, int G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G && flags[i]) contacts[cpos[i]] = i;
}

__global__ void k_bins_fold(double *__restrict__ bins, int nb, const int *__restrict__ nodeBin,
                            const double *__restrict__ nodeF, const int *__restrict__ contacts,
                            const int *__restrict__ countPtr, double *__restrict__ hist2) {
    const int t = threadIdx.x;
    const int count = *countPtr;
    if (t < nb) {
        double a0 = bins[2 * t], a1 = bins[2 * t + 1];
        for (int e = 0; e < count; ++e) {
            const int i = contacts[e];
            if (nodeBin[i] == t) {
                a0 += nodeF[2 * (size_t)i];
                a1 += nodeF[2 * (size_t)i + 1];
            }
        }
        bins[2 * t] = a0;
        bins[2 * t + 1] = a1;
    } else if (t == nb && hist2) {
        double s0 = 0.0, s1 = 0.0;
        for (int e = 0; e < count; ++e) {
            const int i = contacts[e];
            s0 += nodeF[2 * (size_t)i];
            s1 += nodeF[2 * (size_t)i + 1];
        }
        hist2[0] = s0;
        hist2[1] = s1;
    }
}

hipError_t ordered_fold_forces(const KParams &k, OrderedScratch &o, double *hist2,
                               hipStream_t s) {
    hipError_t e;
    const int G = k.gridTotal;
    const unsigned nbk = (unsigned)((G + 1 + 255) / 256);
    hipLaunchKernelGGL(k_contact_flags, dim3(nbk), dim3(256), 0, s, o.nodeBin, o.flags, G);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = exclusive_scan(o.flags, o.cpos, (size_t)G + 1, o.scanTmp, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_contact_compact, dim3(nbk), dim3(256), 0, s, o.flags, o.cpos,
                       o.contacts, G);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int threads = ((k.nbins + 1 + 63) / 64) * 64;
    if (threads > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bins_fold, dim3(1), dim3(threads), 0, s, k.bins, k.nbins, o.nodeBin,
                       o.nodeF, o.contacts, o.cpos + G, hist2);
    return hipGetLastError();
}

}  // namespace icempm
