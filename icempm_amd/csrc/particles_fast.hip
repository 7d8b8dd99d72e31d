// particles_fast.hip -- the particle pass with FMA contraction (as the
// reference's nvcc build has by default): used by FAST mode.  The pragma
// precedes every include so that the inline device math of this translation
// unit is contracted too (it has internal linkage, so the exact build's copy
// is unaffected).
#pragma clang fp contract(fast)

#include <climits>

#include "internal.h"

#define ICEMPM_PASS_NS fast
#define ICEMPM_FAST_ALGEBRA 1
// six wavefronts per SIMD for the wave kernel: the compiler then fits it in
// 80 VGPRs with a 12-byte spill, and the fused pass runs ~3% faster than at
// 82 VGPRs and five wavefronts (measured on the 1.6M-point case)
#define ICEMPM_WAVE_KERNEL_ATTR __attribute__((amdgpu_waves_per_eu(6)))
#include "particles.inc"
