// mlp_device.h -- device pieces of the learned return map shared by the
// training kernels (mlp.hip) and the batch particle pass (batch.hip).
//
// Network: x (5, padded to 16) -> tanh(W1 x + b1) (64) -> tanh(W2 h + b2) (64)
// -> W3 h + b3 (3, padded to 16).  Parameters as one flat vector:
//   W1[64][16] @0, b1[64] @1024, W2[64][64] @1088, b2[64] @5184,
//   W3[16][64] @5248, b3[16] @6272; 6288 values; W[out][in].
// Products run on gfx950's v_mfma_f32_16x16x32_bf16 where the reduction
// length is a multiple of 32 (the 64-wide layers: half the MFMA issues of
// the 16x16x16 form) and on v_mfma_f32_16x16x16_bf16 where it is 16, with
// FP32 accumulation.  Lane layout (wavefront of 64, g = lane/16,
// c = lane%16): lane holds A[c][8g..8g+7] and B[8g..8g+7][c] of a 32-long
// step (A[c][4g..4g+3] and B[4g..4g+3][c] of a 16-long one) and receives
// D[4g+i][c], i < 4.
#pragma once

#include <hip/hip_runtime.h>

namespace icempm {
namespace mlpdev {

constexpr int kIn = 16, kH = 64, kOut = 16;
constexpr int oW1 = 0, oB1 = 1024, oW2 = 1088, oB2 = 5184, oW3 = 5248, oB3 = 6272, kParams = 6288;
constexpr int kRealIn = 5, kRealOut = 3;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// acc += A * B over a 32-long step (operands as BF16 bits)
static __device__ __forceinline__ f4 mfma32(s8 a, s8 b, f4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), acc,
                                                   0, 0, 0);
}

// float -> bf16 bits, round to nearest even (NaN kept quiet)
static __device__ __host__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
static __device__ __host__ __forceinline__ float from_bf16(unsigned short b) {
    return __builtin_bit_cast(float, (unsigned int)b << 16);
}

static __device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One 16x16 output tile of A(16 x K) * B(K x 16) accumulated into acc;
// a(m, k), b(k, n) return BF16 bits.
template <class FA, class FB>
static __device__ __forceinline__ f4 tile_bits(FA a, FB b, int K, f4 acc) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    if (K % 32 == 0) {
        for (int k0 = 0; k0 < K; k0 += 32) {
            s8 av, bv;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                av[j] = (short)a(r, k0 + 8 * g + j);
                bv[j] = (short)b(k0 + 8 * g + j, r);
            }
            acc = mfma32(av, bv, acc);
        }
        return acc;
    }
    for (int k0 = 0; k0 < K; k0 += 16) {
        s4 av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            av[j] = (short)a(r, k0 + 4 * g + j);
            bv[j] = (short)b(k0 + 4 * g + j, r);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// Inference scratch of one wavefront: 16 rows, activations as BF16 bits.
// Rows are padded by two dwords: a lane's four operand values stay one
// aligned 8-byte read, and the 16 rows of a fetch spread over the banks.
constexpr int kInLd = kIn + 4, kHLd = kH + 4;
struct InferScratch {
    alignas(16) unsigned short x[16][kInLd];
    alignas(16) unsigned short h1[16][kHLd];
    alignas(16) unsigned short h2[16][kHLd];
    alignas(16) float y[16][kOut + 1];
};

// One 16x16 output tile of A(16 x K) * W^T, A row-major with stride LD and
// W[n][k] row-major with stride K (both BF16 bits, 8-byte aligned rows):
// each lane's operands are 8-byte reads (two per operand of a 32-long step).
template <int K, int LD>
static __device__ __forceinline__ f4 tile_rows(const unsigned short *A, const unsigned short *W, f4 acc) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    if constexpr (K % 32 == 0) {
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 32) {
            const s4 a0 = *reinterpret_cast<const s4 *>(A + r * LD + k0 + 8 * g);
            const s4 a1 = *reinterpret_cast<const s4 *>(A + r * LD + k0 + 8 * g + 4);
            const s4 b0 = *reinterpret_cast<const s4 *>(W + r * K + k0 + 8 * g);
            const s4 b1 = *reinterpret_cast<const s4 *>(W + r * K + k0 + 8 * g + 4);
            const s8 av = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
            const s8 bv = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
            acc = mfma32(av, bv, acc);
        }
    } else {
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 16) {
            const s4 av = *reinterpret_cast<const s4 *>(A + r * LD + k0 + 4 * g);
            const s4 bv = *reinterpret_cast<const s4 *>(W + r * K + k0 + 4 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
    }
    return acc;
}

// Forward of the 16 rows in s.x.  Wbf: parameters as BF16 bits (weights;
// 8-byte aligned), Pf: the same parameters as FP32 (biases are added in FP32).
static __device__ __forceinline__ void infer16(InferScratch &s, const unsigned short *__restrict__ Wbf,
                                               const float *__restrict__ Pf) {
    static_assert(kInLd % 4 == 0 && kHLd % 4 == 0 && oW1 % 4 == 0 && oW2 % 4 == 0 && oW3 % 4 == 0,
                  "8-byte operand reads");
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
#pragma unroll
    for (int nt = 0; nt < kH / 16; ++nt) {
        const f4 acc = tile_rows<kIn, kInLd>(&s.x[0][0], Wbf + oW1 + nt * 16 * kIn, f4{0, 0, 0, 0});
#pragma unroll
        for (int i = 0; i < 4; ++i) s.h1[4 * g + i][nt * 16 + c] = to_bf16(tanhf(acc[i] + Pf[oB1 + nt * 16 + c]));
    }
    wave_fence();
#pragma unroll
    for (int nt = 0; nt < kH / 16; ++nt) {
        const f4 acc = tile_rows<kH, kHLd>(&s.h1[0][0], Wbf + oW2 + nt * 16 * kH, f4{0, 0, 0, 0});
#pragma unroll
        for (int i = 0; i < 4; ++i) s.h2[4 * g + i][nt * 16 + c] = to_bf16(tanhf(acc[i] + Pf[oB2 + nt * 16 + c]));
    }
    wave_fence();
    const f4 acc = tile_rows<kH, kHLd>(&s.h2[0][0], Wbf + oW3, f4{0, 0, 0, 0});
#pragma unroll
    for (int i = 0; i < 4; ++i) s.y[4 * g + i][c] = acc[i] + Pf[oB3 + c];
    wave_fence();
}

// Normalisation of the map's inputs and outputs (feature = (raw - mean) / scale).
struct Norm {
    float in_mean[5], in_scale[5], out_mean[3], out_scale[3];
};

}  // namespace mlpdev
}  // namespace icempm
