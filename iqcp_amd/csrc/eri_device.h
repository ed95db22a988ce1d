// Rys-quadrature contracted shell-quartet integrals, templated on the angular-momentum
// class (LA LB | LC LD) and the arithmetic type R (double or float).
//
// Algorithm (same integrals as qc-core's shell_eri, integrals/eri/mod.rs:458-854):
//   per primitive quartet: T = rho |P-Q|^2, prefactor 2 pi^(5/2) / (p q sqrt(p+q)) K_ab K_cd
//   (gaussian_product_2e.rs:260-342), Rys nodes/weights (rys.h), per node the B00/B10/B01/C00/C0'
//   coefficients (rys_coefficients.rs:94-180, in the division-free form
//   B00 = u/(2(p+q)), B10 = (1 - u q/(p+q))/(2p), B01 = (1 - u p/(p+q))/(2q),
//   C00 = PA - u q/(p+q) PQ, C0' = QC + u p/(p+q) PQ), 2-D VRR tables per axis
//   (vrr_2d.rs:143-213), then 1-D horizontal transfers per axis (htr_4d.rs:70-354) done ONCE
//   per node for all (i,j,k,l) powers instead of once per Cartesian component, then
//   I(abcd) += Ix Iy Iz summed over nodes and primitive quartets.
#pragma once
#include <utility>
#include "qcg_types.h"
#include "rys.h"

#ifndef QCG_UNROLL_NQ
#define QCG_UNROLL_NQ 128
#endif

namespace qcg {

template <int LA, int LB, int LC, int LD>
struct EriShape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAB = NA * NB, NCD = NC * ND, NQ = NAB * NCD;
    static constexpr int LAB = LA + LB, LCD = LC + LD;
    static constexpr int NROOTS = (LA + LB + LC + LD) / 2 + 1;
    static constexpr int G = (LAB + 1) * (LCD + 1);
    static constexpr int H4 = (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1);
    static constexpr bool kUnroll = NQ <= QCG_UNROLL_NQ;
};

// Component normalisation 1/sqrt((2i-1)!!(2j-1)!!(2k-1)!!) of component c of shell l
// (component-dependent factor of cartesian_gaussian_normalization, integrals/overlap.rs:343-377).
QCG_HD constexpr double comp_norm(int l, int c) {
    double d = dfact_odd(cart_x(l, c)) * dfact_odd(cart_y(l, c)) * dfact_odd(cart_z(l, c));
    return d == 1.0 ? 1.0 : (d == 3.0 ? 0.57735026918962576451 : (d == 9.0 ? 1.0 / 3.0 : 0.0));
}

// The same factors as a literal table: kCompNorm[l][c].
__host__ __device__ constexpr double kCompNorm[3][6] = {
    {1.0, 0, 0, 0, 0, 0},
    {1.0, 1.0, 1.0, 0, 0, 0},
    {0.57735026918962576451, 1.0, 1.0, 0.57735026918962576451, 1.0, 0.57735026918962576451}};

// kCompNorm[l][c] as arithmetic on (l, c): folds to a constant when both are constants after
// unrolling, and needs no memory access when they are not.
QCG_HD double cnorm(int l, int c) {
    return (l == 2 && (c == 0 || c == 3 || c == 5)) ? 0.57735026918962576451 : 1.0;
}

// Per-component indices into the 1-D transfer tables.
template <int LA, int LB, int LC, int LD>
struct CompTab {
    using S = EriShape<LA, LB, LC, LD>;
    int hx[S::NQ], hy[S::NQ], hz[S::NQ];
    constexpr CompTab() : hx(), hy(), hz() {
        for (int a = 0; a < S::NA; ++a)
            for (int b = 0; b < S::NB; ++b)
                for (int c = 0; c < S::NC; ++c)
                    for (int d = 0; d < S::ND; ++d) {
                        const int i = ((a * S::NB + b) * S::NC + c) * S::ND + d;
                        hx[i] = ((cart_x(LA, a) * (LB + 1) + cart_x(LB, b)) * (LC + 1) + cart_x(LC, c)) * (LD + 1) + cart_x(LD, d);
                        hy[i] = ((cart_y(LA, a) * (LB + 1) + cart_y(LB, b)) * (LC + 1) + cart_y(LC, c)) * (LD + 1) + cart_y(LD, d);
                        hz[i] = ((cart_z(LA, a) * (LB + 1) + cart_z(LB, b)) * (LC + 1) + cart_z(LC, c)) * (LD + 1) + cart_z(LD, d);
                    }
    }
};
template <int LA, int LB, int LC, int LD>
inline constexpr CompTab<LA, LB, LC, LD> kCompTab{};

// 2-D VRR table g[n*(NK+1)+m], n <= NB, m <= NK  (build_2d_into, vrr_2d.rs:143-213)
template <int NB, int NK, typename R>
QCG_HD void vrr2d(R* g, R c00, R c0p, R b00, R b10, R b01) {
    constexpr int C = NK + 1;
    g[0] = R(1);
    if constexpr (NB > 0) {
        g[C] = c00;
#pragma unroll
        for (int n = 1; n < NB; ++n) g[(n + 1) * C] = c00 * g[n * C] + R(n) * b10 * g[(n - 1) * C];
    }
#pragma unroll
    for (int m = 0; m < NK; ++m) {
        if (m == 0) g[1] = c0p;
        else g[m + 1] = c0p * g[m] + R(m) * b01 * g[m - 1];
#pragma unroll
        for (int n = 1; n <= NB; ++n) {
            R v = c0p * g[n * C + m] + R(n) * b00 * g[(n - 1) * C + m];
            if (m > 0) v += R(m) * b01 * g[n * C + m - 1];
            g[n * C + m + 1] = v;
        }
    }
}

// 1-D horizontal transfer of one axis: out[((ia*(LB+1)+ib)*(LC+1)+ic)*(LD+1)+id]
// (i,j,k,l+1) = (i,j,k+1,l) + CD (i,j,k,l);  (i,j+1,k,l) = (i+1,j,k,l) + AB (i,j,k,l)
template <int LA, int LB, int LC, int LD, typename R, typename Put>
QCG_HD void hrr1d_put(const R* g, R ab, R cd, Put&& put) {
    constexpr int LAB = LA + LB, LCD = LC + LD, GC = LCD + 1;
    R t[(LAB + 1) * (LCD + 1) * (LD + 1)];
#define T3(n, k, l) t[((n) * (LCD + 1) + (k)) * (LD + 1) + (l)]
#pragma unroll
    for (int n = 0; n <= LAB; ++n) {
#pragma unroll
        for (int k = 0; k <= LCD; ++k) T3(n, k, 0) = g[n * GC + k];
#pragma unroll
        for (int l = 1; l <= LD; ++l)
#pragma unroll
            for (int k = 0; k + l <= LCD; ++k) T3(n, k, l) = T3(n, k + 1, l - 1) + cd * T3(n, k, l - 1);
    }
    R u[(LAB + 1) * (LB + 1)];
#pragma unroll
    for (int k = 0; k <= LC; ++k)
#pragma unroll
        for (int l = 0; l <= LD; ++l) {
#pragma unroll
            for (int i = 0; i <= LAB; ++i) u[i * (LB + 1)] = T3(i, k, l);
#pragma unroll
            for (int j = 1; j <= LB; ++j)
#pragma unroll
                for (int i = 0; i + j <= LAB; ++i)
                    u[i * (LB + 1) + j] = u[(i + 1) * (LB + 1) + j - 1] + ab * u[i * (LB + 1) + j - 1];
#pragma unroll
            for (int i = 0; i <= LA; ++i)
#pragma unroll
                for (int j = 0; j <= LB; ++j) put(((i * (LB + 1) + j) * (LC + 1) + k) * (LD + 1) + l, u[i * (LB + 1) + j]);
        }
#undef T3
}

template <int LA, int LB, int LC, int LD, typename R>
QCG_HD void hrr1d(const R* g, R ab, R cd, R* out) {
    hrr1d_put<LA, LB, LC, LD, R>(g, ab, cd, [&](int h, R v) { out[h] = v; });
}

// body(std::integral_constant<int, I>{}) for I in [0, N): the index is a constant expression in
// the body, so table lookups with it fold at compile time instead of becoming indexed register
// moves or select chains.
template <typename F, int... Is>
QCG_HD void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
QCG_HD void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Run body(i) for i in [0, N): fully unrolled (constant i) when the class is small enough to live
// in registers, a plain loop otherwise (large classes keep their arrays in scratch).
template <bool UNROLL, int N, typename F>
QCG_HD void for_n(F&& body) {
    if constexpr (UNROLL) {
        static_for<N>([&](auto ic) { body((int)decltype(ic)::value); });
    } else {
#pragma unroll 1
        for (int i = 0; i < N; ++i) body(i);
    }
}

// Accumulate the contracted quartet (bra pair P | ket pair Q) into eri[NQ] (row-major
// [a][b][c][d] over Cartesian components; the primitive-dependent normalisation is folded
// into ck, the component normalisation is NOT applied here).  ppb: the bra's primitive pairs
// (P.npp of them, from index 0: a copy staged in LDS or pp + P.pp_off), ppk: the ket's (Q.npp).
// TAB_N: rys_tab already points at the coefficients of this class's Rys order (e.g. a copy
// staged in LDS).
template <int LA, int LB, int LC, int LD, typename R, bool TAB_N = false>
QCG_HD void contract_quartet_pp(const PairRec& P, const PairRec& Q, const PrimPair<R>* __restrict__ ppb,
                                const PrimPair<R>* __restrict__ ppk, const R* __restrict__ rys_tab, R* eri) {
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int N = S::NROOTS;
    const R abx = (R)P.abx, aby = (R)P.aby, abz = (R)P.abz;
    const R cdx = (R)Q.abx, cdy = (R)Q.aby, cdz = (R)Q.abz;
    const R Ax = (R)P.ax, Ay = (R)P.ay, Az = (R)P.az;
    const R Cx = (R)Q.ax, Cy = (R)Q.ay, Cz = (R)Q.az;
    const R two_pi_52 = (R)(2.0 * kPi * kPi * 1.7724538509055160273);
    // ket primitives outer: in the J/K kernels the bra pair is uniform over the block, so the inner
    // loop's bra primitive reads are uniform (an LDS broadcast) and each lane loads its ket
    // primitive once
    for (int j = 0; j < Q.npp; ++j) {
        const PrimPair<R> c = ppk[j];
        const R q = c.p;
        const R hq = c.hp;
        const R QCx = c.px - Cx, QCy = c.py - Cy, QCz = c.pz - Cz;
        for (int i = 0; i < P.npp; ++i) {
            const PrimPair<R> a = ppb[i];
            // the primitive screen k_ab k_cd < 1e-15 (on the FP64 k values) as a factor, not a
            // branch: a screened primitive quartet adds exactly zero, and the loop has no
            // data-dependent exit before its loads (about 5 % of the primitive quartets are screened)
            const R keep = prim_kab(a) * prim_kab(c) < kPrimScreen ? R(0) : R(1);
            const R p = a.p;
            const R hp = a.hp;
            // 1/(p+q) = rs^2: one reciprocal square root per primitive quartet
            const R rs = qrsqrt(p + q);
            const R ipq = rs * rs;
            const R PQx = a.px - c.px, PQy = a.py - c.py, PQz = a.pz - c.pz;
            const R rho = p * q * ipq;
            const R T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
            const R pref = two_pi_52 * rs * a.ckp * c.ckp * keep;
            R rx[N], rw[N];
            if constexpr (TAB_N) rys_eval_tab<R, N>(T, rys_tab, rx, rw);
            else rys_eval<R, N>(T, rys_tab, rx, rw);
            const R qf = q * ipq, pf = p * ipq;
            const R PAx = a.px - Ax, PAy = a.py - Ay, PAz = a.pz - Az;
#pragma unroll
            for (int r = 0; r < N; ++r) {
                const R u = rx[r];
                const R b00 = R(0.5) * u * ipq;
                const R b10 = hp * (R(1) - u * qf);
                const R b01 = hq * (R(1) - u * pf);
                const R uq = u * qf, up = u * pf;
                R gx[S::G], gy[S::G], gz[S::G];
                vrr2d<S::LAB, S::LCD, R>(gx, PAx - uq * PQx, QCx + up * PQx, b00, b10, b01);
                vrr2d<S::LAB, S::LCD, R>(gy, PAy - uq * PQy, QCy + up * PQy, b00, b10, b01);
                vrr2d<S::LAB, S::LCD, R>(gz, PAz - uq * PQz, QCz + up * PQz, b00, b10, b01);
                R ix[S::H4], iy[S::H4], iz[S::H4];
                hrr1d<LA, LB, LC, LD, R>(gx, abx, cdx, ix);
                hrr1d<LA, LB, LC, LD, R>(gy, aby, cdy, iy);
                hrr1d<LA, LB, LC, LD, R>(gz, abz, cdz, iz);
                const R wp = rw[r] * pref;
#pragma unroll
                for (int h = 0; h < S::H4; ++h) iz[h] *= wp;
                if constexpr (S::kUnroll) {
                    static_for<S::NQ>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        constexpr int hx = kCompTab<LA, LB, LC, LD>.hx[k];
                        constexpr int hy = kCompTab<LA, LB, LC, LD>.hy[k];
                        constexpr int hz = kCompTab<LA, LB, LC, LD>.hz[k];
                        eri[k] += ix[hx] * iy[hy] * iz[hz];
                    });
                } else {
                    constexpr const CompTab<LA, LB, LC, LD>& ct = kCompTab<LA, LB, LC, LD>;
                    for_n<false, S::NQ>([&](int k) { eri[k] += ix[ct.hx[k]] * iy[ct.hy[k]] * iz[ct.hz[k]]; });
                }
            }
        }
    }
}

// contract_quartet_pp with both pairs' primitives read from the one array pp.
template <int LA, int LB, int LC, int LD, typename R, bool TAB_N = false>
QCG_HD void contract_quartet(const PairRec& P, const PairRec& Q, const PrimPair<R>* __restrict__ pp,
                             const R* __restrict__ rys_tab, R* eri) {
    contract_quartet_pp<LA, LB, LC, LD, R, TAB_N>(P, Q, pp + P.pp_off, pp + Q.pp_off, rys_tab, eri);
}

}  // namespace qcg
