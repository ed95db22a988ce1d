// Rys quadrature nodes/weights from the piecewise fits in rys_table.inc.
//
// Same quantities as qc-core's rys_roots (crates/qc-core/src/rys/mod.rs:253): n nodes
// u_i = t_i^2 in [0,1) and weights with sum_i w_i u_i^k = F_k(T) for k < 2n.  The reference
// solves for them per call (Boys moments -> Schmidt -> Hessenberg QR, rys/mod.rs:468-583);
// here they are read from fits of the exact rule (see gen_rys_table.py), so every lane
// follows the same instruction stream.  Accuracy of the FP64 fits: ~1e-15 relative.
#pragma once
#include "qcg_common.h"
#include "rys_table.inc"

namespace qcg {

constexpr int kRysStride = QCG_RYS_DEG + 1;

QCG_HD constexpr int rys_tri(int n) { return n * (n - 1) / 2; }

// Number of table entries of the N-point rule (intervals x 2N polynomials x coefficients).
template <int N>
constexpr int rys_tab_size() { return qcg_rys_nint[N] * 2 * N * (QCG_RYS_DEG + 1); }

// c[0] + c[1] s + ... + c[13] s^13 in Estrin's scheme: dependency depth 5 instead of Horner's 13
// (the FP64 FMA latency, not the issue rate, bounds a lone Horner chain per lane).
template <typename R>
QCG_HD R poly13(const R* __restrict__ c, R s) {
    static_assert(QCG_RYS_DEG == 13, "poly13 evaluates the degree-13 fits");
    const R s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
    const R t0 = c[0] + c[1] * s, t1 = c[2] + c[3] * s, t2 = c[4] + c[5] * s, t3 = c[6] + c[7] * s;
    const R t4 = c[8] + c[9] * s, t5 = c[10] + c[11] * s, t6 = c[12] + c[13] * s;
    const R u0 = t0 + t1 * s2, u1 = t2 + t3 * s2, u2 = t4 + t5 * s2;
    const R v0 = u0 + u1 * s4, v1 = u2 + t6 * s4;
    return v0 + v1 * s8;
}

// tabN: the coefficients of the N-point rule, i.e. qcg_rys_coef + qcg_rys_off[N] (or its float
// copy), in device memory or staged in LDS.
template <typename R, int N>
QCG_HD void rys_eval_tab(R T, const R* __restrict__ tabN, R* x, R* w) {
    static_assert(N >= 1 && N <= QCG_RYS_NMAX, "Rys order out of table range");
    if (T >= (R)qcg_rys_tmax[N]) {
        const R rs = qrsqrt(T);
        const R it = rs * rs;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            x[i] = (R)qcg_rys_hx[rys_tri(N) + i] * it;
            w[i] = (R)qcg_rys_hw[rys_tri(N) + i] * rs;
        }
        return;
    }
    const int k = (int)(T * (R)(1.0 / QCG_RYS_H));
    const R s = (T - (R)(k * QCG_RYS_H)) * (R)(2.0 / QCG_RYS_H) - R(1);
    const R* c = tabN + k * (2 * N * kRysStride);
#pragma unroll
    for (int i = 0; i < 2 * N; ++i) {
        const R v = poly13(c + i * kRysStride, s);
        if (i < N) x[i] = v; else w[i - N] = v;
    }
}

// tab: qcg_rys_coef (double) or its float copy, resident in device memory.
template <typename R, int N>
QCG_HD void rys_eval(R T, const R* __restrict__ tab, R* x, R* w) {
    rys_eval_tab<R, N>(T, tab + qcg_rys_off[N], x, w);
}

// Runtime-order variant for host code and batch kernels.
template <typename R>
QCG_HD int rys_eval_n(int n, R T, const R* __restrict__ tab, R* x, R* w) {
    switch (n) {
        case 1: rys_eval<R, 1>(T, tab, x, w); return 0;
        case 2: rys_eval<R, 2>(T, tab, x, w); return 0;
        case 3: rys_eval<R, 3>(T, tab, x, w); return 0;
        case 4: rys_eval<R, 4>(T, tab, x, w); return 0;
        case 5: rys_eval<R, 5>(T, tab, x, w); return 0;
        case 6: rys_eval<R, 6>(T, tab, x, w); return 0;
        case 7: rys_eval<R, 7>(T, tab, x, w); return 0;
        default: return QCG_INVALID_ARG;
    }
}

}  // namespace qcg
