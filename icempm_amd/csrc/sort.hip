// sort.hip -- device-wide exclusive scan, stable LSD radix sort of
// (key, value) int pairs, cell keys and SoA permutation.
//
// Replaces the load-time std::sort of SnapshotManager::SortPoints
// (snapshotmanager.cpp:288-304).  The sort here is stable (ties keep their
// previous order), so the resulting order is a function of the keys alone.
#include <cstdint>

#include "internal.h"

namespace icempm {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// exclusive scan of per-thread values over a 256-thread workgroup
__device__ __forceinline__ int block_exclusive_scan(int v, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        const int s = wsum[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return pre + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const int *in, int *out, size_t n,
                                                            int *blockSums) {
    __shared__ int wsum[kScanThreads / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int total;
    int run = block_exclusive_scan(s, wsum, total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 0 && blockSums) blockSums[blockIdx.x] = total;
}

__global__ void k_scan_add(int *out, size_t n, const int *blockOffsets) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += blockOffsets[i / kScanTile];
}

size_t scan_tmp_ints(size_t n) {
    size_t total = 0;
    size_t m = n;
    while (m > (size_t)kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        total += m + 64;
    }
    return total + 64;
}

hipError_t exclusive_scan(const int *in, int *out, size_t n, int *tmp, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, s, in, out, n,
                           (int *)nullptr);
        return hipGetLastError();
    }
    int *sums = tmp;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n,
                       sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = exclusive_scan(sums, sums, nb, tmp + nb + 64, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n,
                       sums);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Radix sort: 4-bit digits, 2048-key tiles.  Each thread owns 8 consecutive
// keys of its tile, so a tile-wide digit-major scan of per-thread counts
// gives every key a stable rank.
constexpr int kRadixBits = 4;
constexpr int kRadixBuckets = 1 << kRadixBits;

__device__ __forceinline__ unsigned radix_digit(int key, int shift) {
    return (((unsigned)key ^ 0x80000000u) >> shift) & (kRadixBuckets - 1);
}

__global__ __launch_bounds__(kScanThreads) void k_radix_hist(const int *__restrict__ keys,
                                                            size_t n, int shift,
                                                            int *__restrict__ hist,
                                                            int nblocks) {
    __shared__ int cnt[kRadixBuckets];
    if (threadIdx.x < kRadixBuckets) cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kScanTile;
    for (int i = threadIdx.x; i < kScanTile; i += kScanThreads) {
        const size_t g = base + i;
        if (g < n) atomicAdd(&cnt[radix_digit(keys[g], shift)], 1);
    }
    __syncthreads();
    if (threadIdx.x < kRadixBuckets) hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(kScanThreads) void k_radix_scatter(
    const int *__restrict__ keys, const int *__restrict__ vals, int *__restrict__ keysOut,
    int *__restrict__ valsOut, size_t n, int shift, const int *__restrict__ offsets,
    int nblocks) {
    __shared__ int cnt[kRadixBuckets * kScanThreads];   // digit-major [d][t]
    __shared__ int wsum[kScanThreads / 64];
    __shared__ int dstart[kRadixBuckets];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)t * kScanItems;
    int k[kScanItems], v[kScanItems];
    unsigned d[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const bool in = base + i < n;
        k[i] = in ? keys[base + i] : 0;
        v[i] = in ? vals[base + i] : 0;
        d[i] = in ? radix_digit(k[i], shift) : kRadixBuckets;   // sentinel: not stored
    }
#pragma unroll
    for (int b = 0; b < kRadixBuckets; ++b) {
        int c = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) c += (d[i] == (unsigned)b) ? 1 : 0;
        cnt[b * kScanThreads + t] = c;
    }
    __syncthreads();
    // digit-major exclusive scan of the 4096 counts: thread T owns 16
    // consecutive entries
    int mine[kRadixBuckets];
    int s = 0;
#pragma unroll
    for (int i = 0; i < kRadixBuckets; ++i) {
        mine[i] = cnt[t * kRadixBuckets + i];
        s += mine[i];
    }
    int total;
    int run = block_exclusive_scan(s, wsum, total);
#pragma unroll
    for (int i = 0; i < kRadixBuckets; ++i) {
        cnt[t * kRadixBuckets + i] = run;
        run += mine[i];
    }
    __syncthreads();
    if (t < kRadixBuckets) dstart[t] = cnt[t * kScanThreads];
    __syncthreads();
    int seen[kRadixBuckets];
#pragma unroll
    for (int b = 0; b < kRadixBuckets; ++b) seen[b] = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (d[i] < (unsigned)kRadixBuckets) {
            int rank = 0;
#pragma unroll
            for (int b = 0; b < kRadixBuckets; ++b)
                if (d[i] == (unsigned)b) {
                    rank = cnt[b * kScanThreads + t] - dstart[b] + seen[b];
                    seen[b] += 1;
                }
            const int dst = offsets[d[i] * nblocks + blockIdx.x] + rank;
            keysOut[dst] = k[i];
            valsOut[dst] = v[i];
        }
    }
}

__global__ void k_minmax_key(const int *__restrict__ keys, size_t n, unsigned *__restrict__ mm) {
    unsigned lo = 0xffffffffu, hi = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const unsigned u = (unsigned)keys[i] ^ 0x80000000u;
        lo = min(lo, u);
        hi = max(hi, u);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned)__shfl_xor((int)lo, off, 64));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

size_t radix_tmp_ints(size_t n) {
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    const size_t h = nb * kRadixBuckets;
    return 2 * h + scan_tmp_ints(h) + 64;
}

hipError_t radix_sort_pairs(int *keys, int *vals, int *keys2, int *vals2, size_t n, int *tmp,
                            hipStream_t s) {
    if (n <= 1) return hipSuccess;
    const int nb = (int)((n + kScanTile - 1) / kScanTile);
    const size_t h = (size_t)nb * kRadixBuckets;
    int *hist = tmp, *offs = tmp + h, *stmp = tmp + 2 * h;
    unsigned *mm = reinterpret_cast<unsigned *>(stmp + scan_tmp_ints(h));
    // key range decides how many 4-bit passes are needed
    hipError_t e = launch_fill_i32(reinterpret_cast<int *>(mm), 1, -1, s);
    if (e != hipSuccess) return e;
    if ((e = launch_fill_i32(reinterpret_cast<int *>(mm) + 1, 1, 0, s)) != hipSuccess) return e;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_minmax_key, dim3(blocks), dim3(256), 0, s, keys, n, mm);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    unsigned hmm[2];
    if ((e = hipMemcpyAsync(hmm, mm, sizeof(hmm), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    const unsigned diff = hmm[0] ^ hmm[1];
    const int bits = diff ? 32 - __builtin_clz(diff) : 0;
    const int passes = (bits + kRadixBits - 1) / kRadixBits;
    int *kin = keys, *vin = vals, *kout = keys2, *vout = vals2;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kScanThreads), 0, s, kin, n, shift,
                           hist, nb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = exclusive_scan(hist, offs, h, stmp, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kScanThreads), 0, s, kin, vin, kout,
                           vout, n, shift, offs, nb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        int *tk = kin, *tv = vin;
        kin = kout; vin = vout; kout = tk; vout = tv;
    }
    if (kin != keys) {
        if ((e = hipMemcpyAsync(keys, kin, n * sizeof(int), hipMemcpyDeviceToDevice, s)) !=
            hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(vals, vin, n * sizeof(int), hipMemcpyDeviceToDevice, s)) !=
            hipSuccess)
            return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
__global__ void k_point_keys(const double *__restrict__ pts, size_t pp, int n, double hinv,
                             int gridX, int *__restrict__ keys, int *__restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pts[pp * kRowPosX + i], y = pts[pp * (kRowPosX + 1) + i];
    keys[i] = cell_key(x, y, hinv, gridX);
    if (vals) vals[i] = i;
}

hipError_t launch_point_keys(const double *pts, size_t pp, int n, double hinv, int gridX,
                             int *keys, int *vals, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_point_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pts,
                       pp, n, hinv, gridX, keys, vals);
    return hipGetLastError();
}

__global__ void k_raw_keys(const float *__restrict__ xy, int n, double hinv, int gridX,
                           int *__restrict__ keys, int *__restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ax = xy[2 * (size_t)i], ay = xy[2 * (size_t)i + 1];
    keys[i] = (int)(ax * hinv - 0.5) + ((int)(ay * hinv - 0.5)) * gridX;
    vals[i] = i;
}

hipError_t launch_raw_keys(const float *xy, int n, double hinv, int gridX, int *keys, int *vals,
                           hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_raw_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xy, n,
                       hinv, gridX, keys, vals);
    return hipGetLastError();
}

// rows 1..15 are doubles; row 0 holds crushed (u8 at byte i), grain (i16 at
// byte pitch+2i) and the viewer's u8 at byte 3*pitch+i
__global__ void k_permute_points(const double *__restrict__ src, double *__restrict__ dst,
                                 size_t pp, int n, const int *__restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    for (int r = 1; r < kPtsRows; ++r) dst[pp * r + i] = src[pp * r + j];
    const unsigned char *s8 = reinterpret_cast<const unsigned char *>(src);
    unsigned char *d8 = reinterpret_cast<unsigned char *>(dst);
    d8[i] = s8[j];
    reinterpret_cast<int16_t *>(d8 + pp)[i] = reinterpret_cast<const int16_t *>(s8 + pp)[j];
    d8[3 * pp + i] = s8[3 * pp + j];
}

hipError_t launch_permute_points(const double *src, double *dst, size_t pp, int n,
                                 const int *perm, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, pp * kPtsRows * sizeof(double),
                                  hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess || n <= 0) return e;
    hipLaunchKernelGGL(k_permute_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src,
                       dst, pp, n, perm);
    return hipGetLastError();
}

__global__ void k_permute_raw(const float *__restrict__ xy, const int16_t *__restrict__ g,
                              float *__restrict__ xy2, int16_t *__restrict__ g2, int n,
                              const int *__restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    xy2[2 * (size_t)i] = xy[2 * (size_t)j];
    xy2[2 * (size_t)i + 1] = xy[2 * (size_t)j + 1];
    g2[i] = g[j];
}

hipError_t launch_permute_raw(const float *xy, const int16_t *g, float *xy2, int16_t *g2, int n,
                              const int *perm, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_permute_raw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xy, g,
                       xy2, g2, n, perm);
    return hipGetLastError();
}

}  // namespace icempm
