// Order-independent (bitwise reproducible) accumulation of the J and K matrices.
//
// Every contribution v (an FP64 value produced in a fixed order by one lane or one wavefront) is
// split into two integers, h = round(v 2^30) and l = round((v - h 2^-30) 2^61), which are added with
// 64-bit integer atomics into two words per matrix element.  Integer addition is associative, so
// the sums, and the value read back as H 2^-30 + L 2^-61, do not depend on the order in which the
// blocks, waves or lanes of a launch (or of concurrent launches) arrive.  The split is exact for
// |v| < 2^20; a single contribution's rounding is at most 2^-62 (2e-19) absolute, below an FP64
// addition's rounding for any sum larger than 1e-3.  Element totals must stay below 2^32 in
// magnitude (H is an int64 of 2^-30 units); J and K elements of the molecules here are O(10^2).
//
// Layout: element e occupies words [2e] (h) and [2e + 1] (l), so both atomics of one contribution
// touch the same 16-byte line.
#pragma once
#include "qcg_common.h"

namespace qcg {

using fix_t = unsigned long long;

constexpr double kFixMagic = 6755399441055744.0;  // 1.5 * 2^52: adding it rounds to an integer

// h and l of v: fma(v, 2^30, 1.5 2^52) rounds v 2^30 to the nearest integer (|v 2^30| < 2^51) and the
// integer sits in the low mantissa bits, read as the bit pattern's offset from 1.5 2^52.
QCG_HD void fix_split(double v, fix_t& h, fix_t& l) {
    const double m = fma(v, 0x1p30, kFixMagic);
    const double hd = m - kFixMagic;
    const double r = fma(-hd, 0x1p-30, v);  // exact: |r| <= 2^-31
    const double m2 = fma(r, 0x1p61, kFixMagic);
    h = (fix_t)(__builtin_bit_cast(long long, m) - __builtin_bit_cast(long long, kFixMagic));
    l = (fix_t)(__builtin_bit_cast(long long, m2) - __builtin_bit_cast(long long, kFixMagic));
}

QCG_HD double fix_value(fix_t h, fix_t l) {
    return (double)(long long)h * 0x1p-30 + (double)(long long)l * 0x1p-61;
}

#if defined(__HIPCC__)
// acc[e] += v in device memory (agent scope; no return value, so no wait on the atomics).
__device__ __forceinline__ void fix_add_global(fix_t* acc, size_t e, double v) {
    fix_t h, l;
    fix_split(v, h, l);
    __hip_atomic_fetch_add(acc + 2 * e, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(acc + 2 * e + 1, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// acc[e] += v in LDS (workgroup scope)
__device__ __forceinline__ void fix_add_lds(fix_t* acc, int e, double v) {
    fix_t h, l;
    fix_split(v, h, l);
    __hip_atomic_fetch_add(acc + 2 * e, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(acc + 2 * e + 1, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Add an already split pair (h, l) to device memory: the flush of LDS rows, skipping zeros.
__device__ __forceinline__ void fix_add_pair_global(fix_t* acc, size_t e, fix_t h, fix_t l) {
    if (h) __hip_atomic_fetch_add(acc + 2 * e, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (l) __hip_atomic_fetch_add(acc + 2 * e + 1, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double fix_load(const fix_t* acc, size_t e) { return fix_value(acc[2 * e], acc[2 * e + 1]); }
#endif

}  // namespace qcg
