// Shared definitions for the qcg library (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/qcg.h"  // status codes

#define QCG_HD __host__ __device__ __forceinline__

namespace qcg {

constexpr int kMaxL = 2;             // s, p, d (all configured basis sets)
constexpr int kMaxCart = 6;          // (kMaxL+1)(kMaxL+2)/2
constexpr double kPi = 3.14159265358979323846;
constexpr double kSchwarzDefault = 1e-12;   // SCHWARZ_THRESHOLD (integrals/eri/mod.rs:1258)
constexpr double kPrimScreen = 1e-15;       // k_ij, k_ij*k_kl screens (integrals/eri/mod.rs:603,611)

QCG_HD constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
QCG_HD constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }  // la >= lb
constexpr int kNumPairClasses = pair_class(kMaxL, kMaxL) + 1;  // 6

// Cartesian component powers in qc-core order (integrals/cartesian.rs:113-141):
// for i = l..0, for j = l-i..0, k = l-i-j.
QCG_HD constexpr int cart_x(int l, int c) {
    int n = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (n == c) return i;
            ++n;
        }
    return 0;
}
QCG_HD constexpr int cart_y(int l, int c) {
    int n = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (n == c) return j;
            ++n;
        }
    return 0;
}
QCG_HD constexpr int cart_z(int l, int c) { return l - cart_x(l, c) - cart_y(l, c); }

// Reciprocal and reciprocal square root in the arithmetic type R, for positive normal arguments
// (exponent sums, T above the asymptotic switch).  On the device:
//   FP32: the hardware v_rcp_f32 / v_rsq_f32 (about 1 ulp), all the FP32 quartets of the mixed
//         mode need (their contributions are below 1e-6 x 1e-7);
//   FP64: the hardware v_rcp_f64 / v_rsq_f64 estimate (5e-8 relative, measured on 2e5 arguments
//         over 1e-8..1e8) refined by one Newton step (r <- r + r(1 - x r);
//         y <- y + y/2 (1 - x y^2)): 3e-15 / 4e-15 relative, far below what the integrals need
//         (1e-12 relative keeps energies to 1e-9 Ha), at a fraction of IEEE division / sqrt.
// On the host: plain division / sqrt.
template <typename R>
QCG_HD R qrcp(R x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(R) == 4) {
        return __builtin_amdgcn_rcpf(x);
    } else {
        // v_rcp_f64 is good to 5e-8 relative (measured); one Newton step gives 3e-15
        const R r = __builtin_amdgcn_rcp(x);
        const R e = fma(-x, r, R(1));
        return fma(r, e, r);
    }
#else
    return R(1) / x;
#endif
}
template <typename R>
QCG_HD R qrsqrt(R x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(R) == 4) {
        return __builtin_amdgcn_rsqf(x);
    } else {
        // v_rsq_f64 is good to 5e-8 relative (measured); one Newton step gives 4e-15
        const R y = __builtin_amdgcn_rsq(x);
        const R e = fma(-x * y, y, R(1));
        return fma(R(0.5) * y, e, y);
    }
#else
    return R(1) / sqrt(x);
#endif
}

QCG_HD constexpr double dfact_odd(int n) {  // (2n-1)!! with (-1)!! = 1
    double r = 1.0;
    for (int k = 2 * n - 1; k > 1; k -= 2) r *= k;
    return r;
}

}  // namespace qcg
