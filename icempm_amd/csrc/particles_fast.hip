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
#include "particles.inc"
