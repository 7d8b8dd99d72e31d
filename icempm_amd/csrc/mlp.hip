// mlp.hip -- the learned return map: a 5 -> 64 -> 64 -> 3 MLP with tanh
// hidden layers, BF16 weights and activations on the MFMA units
// (v_mfma_f32_16x16x32_bf16 for the 64-long products, v_mfma_f32_16x16x16_bf16
// for the 16-long ones, FP32 accumulation), trained on the GPU by the
// kernels below (forward, backward, Adam) and evaluated by k_mlp_infer and by
// the batch particle pass.
//
// Parameters (one flat FP32 vector, also kept as BF16 for the MFMA operands):
//   W1[64][16]  (inputs padded 5 -> 16 with zero columns)   offset 0
//   b1[64]                                                  offset 1024
//   W2[64][64]                                              offset 1088
//   b2[64]                                                  offset 5184
//   W3[16][64]  (outputs padded 3 -> 16 with zero rows)     offset 5248
//   b3[16]                                                  offset 6272
// W[out][in]: y = W x + b.  Inputs and outputs are normalised by the caller.
//
// MFMA lane layout (wavefront of 64, g = l/16): a 32-long step gives lane l
// A[l%16][8g..8g+7] and B[8g..8g+7][l%16], a 16-long step A[l%16][4g..4g+3]
// and B[4g..4g+3][l%16]; lane l receives D[4g+i][l%16], i < 4.
// tests/test_gpu_mlp.py checks both layouts against a CPU product.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/icempm.h"
#include "internal.h"
#include "mlp_device.h"

namespace icempm {
namespace mlp {

using namespace mlpdev;

// One 16x16 output tile of A(16 x K) * B(K x 16) accumulated into acc.
// a(m, k) and b(k, n) return the operands as floats (rounded to BF16 here).
template <class FA, class FB>
static __device__ __forceinline__ f4 tile(FA a, FB b, int K, f4 acc) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    if (K % 32 == 0) {
        for (int k0 = 0; k0 < K; k0 += 32) {
            s8 av, bv;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                av[j] = (short)to_bf16(a(r, k0 + 8 * g + j));
                bv[j] = (short)to_bf16(b(k0 + 8 * g + j, r));
            }
            acc = mfma32(av, bv, acc);
        }
        return acc;
    }
    for (int k0 = 0; k0 < K; k0 += 16) {
        s4 av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            av[j] = (short)to_bf16(a(r, k0 + 4 * g + j));
            bv[j] = (short)to_bf16(b(k0 + 4 * g + j, r));
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

// Lane layout probe: D = A * B, A (16 x K) and B (K x 16) row-major, by the
// tile routine of the training kernels (K = 16: the 16x16x16 form; K a
// multiple of 32: the 16x16x32 form).
__global__ void k_mfma_probe(const float *A, const float *B, float *D, int K) {
    const int l = threadIdx.x, g = l >> 4, c = l & 15;
    f4 acc = {0, 0, 0, 0};
    acc = tile([&](int m, int k) { return A[m * K + k]; }, [&](int k, int n) { return B[k * 16 + n]; },
               K, acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) D[(4 * g + i) * 16 + c] = acc[i];
}

// Per-wavefront scratch in shared memory (floats).
struct WaveScratch {
    float x[16][kIn];
    float h1[16][kH];
    float h2[16][kH];
    float y[16][kOut];
    float d2[16][kH];   // dL/d(pre-activation 2)
    float d1[16][kH];   // dL/d(pre-activation 1)
};

// store a D tile (lane l: rows 4g+i, column c) into a row-major array
template <int LD>
static __device__ __forceinline__ void put(float (*dst)[LD], int n0, f4 v) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[4 * g + i][n0 + c] = v[i];
}

// forward of one 16-row tile held in s.x; parameters P (float, shared or global)
template <class PW>
static __device__ __forceinline__ void forward16(WaveScratch &s, const PW &P) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    for (int nt = 0; nt < kH / 16; ++nt) {
        f4 acc = {0, 0, 0, 0};
        acc = tile([&](int m, int k) { return s.x[m][k]; },
                   [&](int k, int n) { return P[oW1 + (nt * 16 + n) * kIn + k]; }, kIn, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = tanhf(acc[i] + P[oB1 + nt * 16 + c]);
        put<kH>(s.h1, nt * 16, acc);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int nt = 0; nt < kH / 16; ++nt) {
        f4 acc = {0, 0, 0, 0};
        acc = tile([&](int m, int k) { return s.h1[m][k]; },
                   [&](int k, int n) { return P[oW2 + (nt * 16 + n) * kH + k]; }, kH, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = tanhf(acc[i] + P[oB2 + nt * 16 + c]);
        put<kH>(s.h2, nt * 16, acc);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    {
        f4 acc = {0, 0, 0, 0};
        acc = tile([&](int m, int k) { return s.h2[m][k]; },
                   [&](int k, int n) { return P[oW3 + n * kH + k]; }, kH, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] += P[oB3 + c];
        put<kOut>(s.y, 0, acc);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    (void)g;
}

// Inference: rows of normalised inputs (n x 5) -> normalised outputs (n x 3),
// by the same wavefront routine as the learned return map of the batch pass.
__global__ __launch_bounds__(256) void k_mlp_infer(const unsigned short *__restrict__ wbf,
                                                   const float *__restrict__ params,
                                                   const float *__restrict__ X, int n,
                                                   float *__restrict__ Y) {
    __shared__ InferScratch ws[4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    InferScratch &s = ws[w];
    const int tiles = (n + 15) / 16;
    for (int t = blockIdx.x * 4 + w; t < tiles; t += gridDim.x * 4) {
        for (int e = l; e < 16 * kIn; e += 64) {
            const int m = e / kIn, k = e % kIn, row = t * 16 + m;
            s.x[m][k] = to_bf16((row < n && k < kRealIn) ? X[(size_t)row * kRealIn + k] : 0.f);
        }
        wave_fence();
        infer16(s, wbf, params);
        for (int e = l; e < 16 * kRealOut; e += 64) {
            const int m = e / kRealOut, k = e % kRealOut, row = t * 16 + m;
            if (row < n) Y[(size_t)row * kRealOut + k] = s.y[m][k];
        }
        wave_fence();
    }
}

// Gradient of L = 1/(2B) sum_rows sum_{k<3} (y_k - t_k)^2 over rows
// [r0, r0 + B): every wavefront accumulates its tiles' gradient in registers
// and writes one partial vector (kParams floats) to part[wave].  Also adds
// the tiles' loss to *loss.
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ params,
                                                  const float *__restrict__ X,
                                                  const float *__restrict__ T, int r0, int B,
                                                  float *__restrict__ part, float *__restrict__ loss) {
    __shared__ WaveScratch ws[4];
    __shared__ float P[kParams];
    for (int e = threadIdx.x; e < kParams; e += 256) P[e] = params[e];
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    WaveScratch &s = ws[w];
    const float invB = 1.f / (float)B;
    f4 gW1[4], gW2[16], gW3[4];
    float gb1[4], gb2[4], gb3 = 0;
    for (int i = 0; i < 4; ++i) {
        gW1[i] = f4{0, 0, 0, 0};
        gW3[i] = f4{0, 0, 0, 0};
        gb1[i] = 0;
        gb2[i] = 0;
    }
    for (int i = 0; i < 16; ++i) gW2[i] = f4{0, 0, 0, 0};
    float lsum = 0;
    const int tiles = (B + 15) / 16;
    const int wave = blockIdx.x * 4 + w, nwave = gridDim.x * 4;
    for (int t = wave; t < tiles; t += nwave) {
        for (int e = l; e < 16 * kIn; e += 64) {
            const int m = e / kIn, k = e % kIn, row = t * 16 + m;
            s.x[m][k] = (row < B && k < kRealIn) ? X[(size_t)(r0 + row) * kRealIn + k] : 0.f;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        forward16(s, P);
        // dY (kept in s.y), zero for padded outputs and rows
        for (int e = l; e < 16 * kOut; e += 64) {
            const int m = e / kOut, k = e % kOut, row = t * 16 + m;
            float d = 0;
            if (row < B && k < kRealOut) {
                d = s.y[m][k] - T[(size_t)(r0 + row) * kRealOut + k];
                lsum += 0.5f * d * d;
                d *= invB;
            }
            s.y[m][k] = d;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // layer 3: gW3[o][h] += sum_m dY[m][o] h2[m][h]; gb3[o] += sum_m dY[m][o]
        for (int ht = 0; ht < 4; ++ht)
            gW3[ht] = tile([&](int o, int m) { return s.y[m][o]; },
                           [&](int m, int h) { return s.h2[m][ht * 16 + h]; }, 16, gW3[ht]);
        // d2 = (dY W3) * (1 - h2^2)
        for (int ht = 0; ht < 4; ++ht) {
            f4 acc = {0, 0, 0, 0};
            acc = tile([&](int m, int o) { return s.y[m][o]; },
                       [&](int o, int h) { return P[oW3 + o * kH + ht * 16 + h]; }, 16, acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float hv = s.h2[4 * g + i][ht * 16 + c];
                acc[i] *= (1.f - hv * hv);
            }
            put<kH>(s.d2, ht * 16, acc);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // layer 2: gW2[o][h] += sum_m d2[m][o] h1[m][h]
        for (int ot = 0; ot < 4; ++ot)
            for (int ht = 0; ht < 4; ++ht)
                gW2[ot * 4 + ht] = tile([&](int o, int m) { return s.d2[m][ot * 16 + o]; },
                                        [&](int m, int h) { return s.h1[m][ht * 16 + h]; }, 16,
                                        gW2[ot * 4 + ht]);
        // d1 = (d2 W2) * (1 - h1^2)
        for (int ht = 0; ht < 4; ++ht) {
            f4 acc = {0, 0, 0, 0};
            acc = tile([&](int m, int o) { return s.d2[m][o]; },
                       [&](int o, int h) { return P[oW2 + o * kH + ht * 16 + h]; }, kH, acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float hv = s.h1[4 * g + i][ht * 16 + c];
                acc[i] *= (1.f - hv * hv);
            }
            put<kH>(s.d1, ht * 16, acc);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // layer 1: gW1[o][k] += sum_m d1[m][o] x[m][k]
        for (int ot = 0; ot < 4; ++ot)
            gW1[ot] = tile([&](int o, int m) { return s.d1[m][ot * 16 + o]; },
                           [&](int m, int k) { return s.x[m][k]; }, 16, gW1[ot]);
        // biases: column sums (lanes with g == 0 keep them)
        for (int i = 0; i < 4; ++i) {
            float a = 0, b = 0;
            for (int m = g; m < 16; m += 4) {
                a += s.d1[m][i * 16 + c];
                b += s.d2[m][i * 16 + c];
            }
            a += __shfl_xor(a, 16, 64);
            a += __shfl_xor(a, 32, 64);
            b += __shfl_xor(b, 16, 64);
            b += __shfl_xor(b, 32, 64);
            gb1[i] += a;
            gb2[i] += b;
        }
        {
            float a = 0;
            for (int m = g; m < 16; m += 4) a += s.y[m][c];
            a += __shfl_xor(a, 16, 64);
            a += __shfl_xor(a, 32, 64);
            gb3 += a;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    // write this wavefront's partial gradient
    float *out = part + (size_t)wave * kParams;
    for (int ot = 0; ot < 4; ++ot)
        for (int i = 0; i < 4; ++i) out[oW1 + (ot * 16 + 4 * g + i) * kIn + c] = gW1[ot][i];
    for (int ot = 0; ot < 4; ++ot)
        for (int ht = 0; ht < 4; ++ht)
            for (int i = 0; i < 4; ++i)
                out[oW2 + (ot * 16 + 4 * g + i) * kH + ht * 16 + c] = gW2[ot * 4 + ht][i];
    for (int ht = 0; ht < 4; ++ht)
        for (int i = 0; i < 4; ++i) out[oW3 + (4 * g + i) * kH + ht * 16 + c] = gW3[ht][i];
    if (g == 0) {
        for (int i = 0; i < 4; ++i) {
            out[oB1 + i * 16 + c] = gb1[i];
            out[oB2 + i * 16 + c] = gb2[i];
        }
        out[oB3 + c] = gb3;
    }
    // the loss of this wavefront's rows
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
    if (l == 0) atomicAdd(loss, lsum);
}

// grad = sum over the partial vectors; then one Adam step on the FP32
// master parameters.  Parameters that are padding stay zero.
// The BF16 copy used by inference is refreshed from the result.
// A workgroup of 256 handles 64 parameters: its four wavefronts each sum a
// quarter of the partial vectors (64 consecutive floats per load), then the
// first wavefront adds the four sums and takes the step.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ params, unsigned short *__restrict__ wbf,
                                                  float *__restrict__ mom1, float *__restrict__ mom2,
                                                  const float *__restrict__ part, int nparts,
                                                  float *__restrict__ grad, float lr, float b1, float b2,
                                                  float eps, float c1, float c2) {
    __shared__ float qs[4][64];
    const int l = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + l;
    float g4 = 0;
    if (i < kParams)
        for (int p = q; p < nparts; p += 4) g4 += part[(size_t)p * kParams + i];
    qs[q][l] = g4;
    __syncthreads();
    if (q != 0 || i >= kParams) return;
    const float gsum = (qs[0][l] + qs[1][l]) + (qs[2][l] + qs[3][l]);
    grad[i] = gsum;
    // padding: input columns >= 5 of W1, output rows >= 3 of W3 and b3
    bool pad = false;
    if (i < oB1) pad = (i % kIn) >= kRealIn;
    else if (i >= oW3 && i < oB3) pad = ((i - oW3) / kH) >= kRealOut;
    else if (i >= oB3) pad = (i - oB3) >= kRealOut;
    if (pad) {
        params[i] = 0;
        wbf[i] = 0;
        return;
    }
    if (lr != 0) {
        const float m = b1 * mom1[i] + (1 - b1) * gsum;
        const float v = b2 * mom2[i] + (1 - b2) * gsum * gsum;
        mom1[i] = m;
        mom2[i] = v;
        params[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
    wbf[i] = to_bf16(params[i]);
}

}  // namespace mlp
}  // namespace icempm

using namespace icempm::mlp;

struct icempm_mlp {
    hipStream_t stream = nullptr;
    float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *part = nullptr;
    unsigned short *wbf = nullptr;   // params as BF16 bits
    float *loss = nullptr;
    int nparts = 0;
    int t = 0;   // Adam step
    std::string lastError;
};

namespace {
constexpr int kGradBlocks = 128;

int mfail(icempm_mlp *m, hipError_t e, const char *tag) {
    if (m) m->lastError = std::string(tag) + ": " + hipGetErrorString(e);
    return ICEMPM_E_HIP;
}
#define MCHK(m, call, tag)                               \
    do {                                                 \
        hipError_t _e = (call);                          \
        if (_e != hipSuccess) return mfail(m, _e, tag);  \
    } while (0)
}  // namespace

extern "C" {

int icempm_mlp_num_params(void) { return kParams; }

int icempm_mlp_create(icempm_mlp **out, const float *init) {
    if (!out || !init) return ICEMPM_E_ARG;
    icempm_mlp *m = new icempm_mlp();
    m->nparts = kGradBlocks * 4;
    hipError_t e = hipSetDevice(0);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&m->params, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->m1, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->m2, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->grad, kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->part, (size_t)m->nparts * kParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->loss, sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->wbf, kParams * sizeof(unsigned short));
    if (e == hipSuccess) e = hipMemcpy(m->params, init, kParams * sizeof(float), hipMemcpyHostToDevice);
    std::vector<unsigned short> wb(kParams);
    for (int i = 0; i < kParams; ++i) wb[i] = to_bf16(init[i]);
    if (e == hipSuccess) e = hipMemcpy(m->wbf, wb.data(), kParams * sizeof(unsigned short), hipMemcpyHostToDevice);
    std::vector<float> z(kParams, 0.f);
    if (e == hipSuccess) e = hipMemcpy(m->m1, z.data(), kParams * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->m2, z.data(), kParams * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        icempm_mlp_destroy(m);
        return ICEMPM_E_HIP;
    }
    *out = m;
    return ICEMPM_OK;
}

void icempm_mlp_destroy(icempm_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (float *p : {m->params, m->m1, m->m2, m->grad, m->part, m->loss})
        if (p) (void)hipFree(p);
    if (m->wbf) (void)hipFree(m->wbf);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char *icempm_mlp_last_error(const icempm_mlp *m) { return m ? m->lastError.c_str() : ""; }

int icempm_mlp_get_params(icempm_mlp *m, float *out) {
    if (!m || !out) return ICEMPM_E_ARG;
    MCHK(m, hipMemcpy(out, m->params, kParams * sizeof(float), hipMemcpyDeviceToHost), "mlp get");
    return ICEMPM_OK;
}

const float *icempm_mlp_device_params(const icempm_mlp *m) { return m ? m->params : nullptr; }

// X: n x 5 normalised inputs, T: n x 3 normalised targets (device pointers).
// Gradient over rows [r0, r0+B) into *grad_out (host, may be NULL) and the
// mean loss; lr > 0 also takes one Adam step.
int icempm_mlp_step(icempm_mlp *m, const float *X, const float *T, int r0, int B, float lr,
                    float *grad_out, float *loss_out) {
    if (!m || !X || !T || B <= 0 || r0 < 0) return ICEMPM_E_ARG;
    MCHK(m, icempm::launch_fill_i32(reinterpret_cast<int *>(m->loss), 1, 0, m->stream), "mlp step");
    hipLaunchKernelGGL(k_mlp_grad, dim3(kGradBlocks), dim3(256), 0, m->stream, m->params, X, T, r0, B,
                       m->part, m->loss);
    MCHK(m, hipGetLastError(), "mlp grad");
    const float b1 = 0.9f, b2 = 0.999f;
    if (lr > 0) m->t += 1;
    const float c1 = 1 - std::pow(b1, (float)(m->t > 0 ? m->t : 1));
    const float c2 = 1 - std::pow(b2, (float)(m->t > 0 ? m->t : 1));
    hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 63) / 64), dim3(256), 0, m->stream, m->params, m->wbf,
                       m->m1, m->m2, m->part, m->nparts, m->grad, lr, b1, b2, 1e-8f, c1, c2);
    MCHK(m, hipGetLastError(), "mlp adam");
    if (grad_out)
        MCHK(m, hipMemcpyAsync(grad_out, m->grad, kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream),
             "mlp step");
    float l = 0;
    MCHK(m, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream), "mlp step");
    MCHK(m, hipStreamSynchronize(m->stream), "mlp step");
    if (loss_out) *loss_out = l / (float)B;
    return ICEMPM_OK;
}

// Y (device, n x 3) = MLP(X (device, n x 5))
int icempm_mlp_infer(icempm_mlp *m, const float *X, int n, float *Y) {
    if (!m || !X || !Y || n < 0) return ICEMPM_E_ARG;
    if (n == 0) return ICEMPM_OK;
    const int tiles = (n + 15) / 16;
    int blocks = (tiles + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, m->stream, m->wbf, m->params, X, n, Y);
    MCHK(m, hipGetLastError(), "mlp infer");
    MCHK(m, hipStreamSynchronize(m->stream), "mlp infer");
    return ICEMPM_OK;
}

int icempm_device_alloc(size_t nbytes, void **ptr) {
    if (!ptr) return ICEMPM_E_ARG;
    *ptr = nullptr;
    return hipMalloc(ptr, nbytes < 1 ? 1 : nbytes) == hipSuccess ? ICEMPM_OK : ICEMPM_E_HIP;
}

void icempm_device_free(void *ptr) {
    if (ptr) (void)hipFree(ptr);
}

int icempm_copy_to_device(void *dst, const void *src, size_t nbytes) {
    if (!dst || !src) return ICEMPM_E_ARG;
    return hipMemcpy(dst, src, nbytes, hipMemcpyHostToDevice) == hipSuccess ? ICEMPM_OK : ICEMPM_E_HIP;
}

int icempm_copy_to_host(void *dst, const void *src, size_t nbytes) {
    if (!dst || !src) return ICEMPM_E_ARG;
    return hipMemcpy(dst, src, nbytes, hipMemcpyDeviceToHost) == hipSuccess ? ICEMPM_OK : ICEMPM_E_HIP;
}

// The training driver (what icempm_amd/learned.py's train_return_map did in
// Python): normalisation fitted on the pairs, one shuffle of the rows, Glorot
// initial weights, `steps` Adam steps on consecutive windows of `batch` rows
// with a cosine learning-rate decay to lr / 20, and the loss over all rows.
namespace {

unsigned long long splitmix64(unsigned long long &s) {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(unsigned long long &s) { return ((splitmix64(s) >> 11) + 0.5) * 0x1.0p-53; }
// a standard normal variate (Box-Muller)
double normal01(unsigned long long &s) {
    const double u = uniform01(s), v = uniform01(s);
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
}

}  // namespace

int icempm_mlp_train(const float *X, const float *T, int n, int steps, int batch, float lr,
                     unsigned long long seed, float *params_out, float *norm_out, float *trace,
                     int trace_every, float *final_loss) {
    if (!X || !T || n <= 0 || steps < 0 || batch <= 0 || !(lr >= 0.f) || !params_out || !norm_out ||
        (trace && trace_every <= 0))
        return ICEMPM_E_ARG;
    // normalisation: mean and (population) standard deviation per column;
    // a constant input column (the hardening feature) gets mean 0, scale 1
    double im[kRealIn] = {}, is[kRealIn] = {}, om[kRealOut] = {}, os[kRealOut] = {};
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < kRealIn; ++k) im[k] += X[(size_t)i * kRealIn + k];
        for (int k = 0; k < kRealOut; ++k) om[k] += T[(size_t)i * kRealOut + k];
    }
    for (int k = 0; k < kRealIn; ++k) im[k] /= n;
    for (int k = 0; k < kRealOut; ++k) om[k] /= n;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < kRealIn; ++k) {
            const double d = X[(size_t)i * kRealIn + k] - im[k];
            is[k] += d * d;
        }
        for (int k = 0; k < kRealOut; ++k) {
            const double d = T[(size_t)i * kRealOut + k] - om[k];
            os[k] += d * d;
        }
    }
    for (int k = 0; k < kRealIn; ++k) {
        is[k] = std::sqrt(is[k] / n);
        if (is[k] < 1e-12) {
            im[k] = 0.0;
            is[k] = 1.0;
        }
    }
    for (int k = 0; k < kRealOut; ++k) {
        os[k] = std::sqrt(os[k] / n);
        if (os[k] < 1e-30) os[k] = 1.0;
    }
    for (int k = 0; k < kRealIn; ++k) {
        norm_out[k] = (float)im[k];
        norm_out[kRealIn + k] = (float)is[k];
    }
    for (int k = 0; k < kRealOut; ++k) {
        norm_out[2 * kRealIn + k] = (float)om[k];
        norm_out[2 * kRealIn + kRealOut + k] = (float)os[k];
    }
    // the rows shuffled once (Fisher-Yates), normalised
    unsigned long long rs = seed;
    std::vector<int> order((size_t)n);
    for (int i = 0; i < n; ++i) order[(size_t)i] = i;
    for (int i = n - 1; i > 0; --i) {
        const int j = (int)(splitmix64(rs) % (unsigned long long)(i + 1));
        std::swap(order[(size_t)i], order[(size_t)j]);
    }
    std::vector<float> Xn((size_t)n * kRealIn), Tn((size_t)n * kRealOut);
    for (int r = 0; r < n; ++r) {
        const size_t i = (size_t)order[(size_t)r];
        for (int k = 0; k < kRealIn; ++k) Xn[(size_t)r * kRealIn + k] = (float)((X[i * kRealIn + k] - im[k]) / is[k]);
        for (int k = 0; k < kRealOut; ++k)
            Tn[(size_t)r * kRealOut + k] = (float)((T[i * kRealOut + k] - om[k]) / os[k]);
    }
    // Glorot-scaled normal initial weights, zero biases and padding
    std::vector<float> P(kParams, 0.f);
    for (int o = 0; o < kH; ++o)
        for (int k = 0; k < kRealIn; ++k) P[oW1 + o * kIn + k] = (float)(normal01(rs) * std::sqrt(2.0 / (kRealIn + kH)));
    for (int e = 0; e < kH * kH; ++e) P[oW2 + e] = (float)(normal01(rs) * std::sqrt(2.0 / (2 * kH)));
    for (int o = 0; o < kRealOut; ++o)
        for (int k = 0; k < kH; ++k) P[oW3 + o * kH + k] = (float)(normal01(rs) * std::sqrt(2.0 / (kRealOut + kH)));
    float *dX = nullptr, *dT = nullptr;
    icempm_mlp *m = nullptr;
    int rc = ICEMPM_OK;
    if (icempm_device_alloc(Xn.size() * sizeof(float), reinterpret_cast<void **>(&dX)) != ICEMPM_OK ||
        icempm_device_alloc(Tn.size() * sizeof(float), reinterpret_cast<void **>(&dT)) != ICEMPM_OK ||
        icempm_copy_to_device(dX, Xn.data(), Xn.size() * sizeof(float)) != ICEMPM_OK ||
        icempm_copy_to_device(dT, Tn.data(), Tn.size() * sizeof(float)) != ICEMPM_OK)
        rc = ICEMPM_E_HIP;
    if (rc == ICEMPM_OK) rc = icempm_mlp_create(&m, P.data());
    if (rc == ICEMPM_OK) {
        const int B = batch < n ? batch : n;
        const int windows = n / B > 1 ? n / B : 1;
        for (int it = 0; it < steps && rc == ICEMPM_OK; ++it) {
            const int r0 = (it % windows) * B;
            const double a =
                (double)lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * it / (steps > 0 ? steps : 1))));
            float l = 0;
            rc = icempm_mlp_step(m, dX, dT, r0, B, (float)a, nullptr, &l);
            if (rc == ICEMPM_OK && trace && it % trace_every == 0) trace[it / trace_every] = l;
        }
        float full = 0;
        if (rc == ICEMPM_OK) rc = icempm_mlp_step(m, dX, dT, 0, n, 0.f, nullptr, &full);
        if (rc == ICEMPM_OK && final_loss) *final_loss = full;
        if (rc == ICEMPM_OK) rc = icempm_mlp_get_params(m, params_out);
    }
    if (m) icempm_mlp_destroy(m);
    icempm_device_free(dX);
    icempm_device_free(dT);
    return rc;
}

// D (16 x 16, host) = A (16 x K) * B (K x 16) through the MFMA tile routine,
// for the lane-layout test; K is 16, 32 or 64.
int icempm_mfma_probe_k(const float *A, const float *B, float *D, int K) {
    if (!A || !B || !D || (K != 16 && K != 32 && K != 64)) return ICEMPM_E_ARG;
    const size_t na = 16 * (size_t)K, nb = (size_t)K * 16;
    float *d = nullptr;
    hipError_t e = hipMalloc(&d, (na + nb + 256) * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d, A, na * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + na, B, nb * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, d, d + na, d + na + nb, K);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(D, d + na + nb, 256 * sizeof(float), hipMemcpyDeviceToHost);
    if (d) (void)hipFree(d);
    return e == hipSuccess ? ICEMPM_OK : ICEMPM_E_HIP;
}

int icempm_mfma_probe(const float *A, const float *B, float *D) { return icempm_mfma_probe_k(A, B, D, 16); }

}  // extern "C"
