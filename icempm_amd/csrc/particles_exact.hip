// particles_exact.hip -- the particle pass without FMA contraction (the
// build's -ffp-contract=off): used by ORDERED mode, where the results must
// equal a sequential restatement of the reference bit for bit.
#include <climits>

#include "internal.h"

#define ICEMPM_PASS_NS exact
#include "particles.inc"
