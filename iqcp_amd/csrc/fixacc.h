// Order-independent (bitwise reproducible) accumulation of the J and K matrices.
//
// Every contribution v (an FP64 value produced in a fixed order by one lane or one wavefront) is
// rounded to an integer multiple of 2^-S, w = round(v 2^S), and added with a 64-bit integer atomic.
// Integer addition is associative (two's-complement wrap-around included), so the sums, and the
// value read back as W 2^-S, do not depend on the order in which the blocks, waves or lanes of a
// launch, or of concurrent launches, arrive.
//
// S is chosen per build and molecule from a bound B on the accumulated magnitudes (fix_scale_kernel,
// eri_misc.hip): the largest S with 2^S B < 2^61, separately for J and K.  For the molecules here B is
// 2^9..2^11, so S is 50..52 and one contribution rounds by at most 2^-51..2^-53 (1e-16 absolute),
// which is what one FP64 addition into a sum of magnitude 1 rounds by.
#pragma once
#include "qcg_common.h"

namespace qcg {

using fix_t = unsigned long long;

// round(v * scale) as a two's-complement 64-bit word (|v * scale| < 2^62)
QCG_HD fix_t fix_word(double v, double scale) { return (fix_t)(long long)rint(v * scale); }

// the accumulated value of a word (inv_scale = 2^-S)
QCG_HD double fix_value(fix_t w, double inv_scale) { return (double)(long long)w * inv_scale; }

// 2^S for a bound B on the magnitudes: the largest S <= 60 with 2^S B < 2^61 (S >= 0)
QCG_HD double fix_scale_for(double B) {
    int e = 0;
    (void)frexp(B > 0.0 ? B : 1.0, &e);  // B < 2^e
    int S = 61 - e;
    S = S > 60 ? 60 : (S < 0 ? 0 : S);
    return ldexp(1.0, S);
}

#if defined(__HIPCC__)
// acc[e] += v in device memory (agent scope; no return value, so no wait on the atomic)
__device__ __forceinline__ void fix_add_global(fix_t* acc, size_t e, double v, double scale) {
    __hip_atomic_fetch_add(acc + e, fix_word(v, scale), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// acc[e] += v in LDS (workgroup scope)
__device__ __forceinline__ void fix_add_lds(fix_t* acc, int e, double v, double scale) {
    __hip_atomic_fetch_add(acc + e, fix_word(v, scale), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Add an already converted word to device memory (the flush of LDS rows), skipping zeros.
__device__ __forceinline__ void fix_add_word_global(fix_t* acc, size_t e, fix_t w) {
    if (w) __hip_atomic_fetch_add(acc + e, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
#endif

}  // namespace qcg
