// Per-quartet adjoint contraction of the derivative ERIs (host/device), used by eri_grad.hip.
// gq[0..8] = sum_abcd W_abcd d(ab|cd)/d(A, B, C)_xyz for one contracted quartet; the density
// weights W (norms and symmetry factor folded in) come from the caller.
#pragma once
#include "eri_device.h"

namespace qcg {

template <int LA, int LB, int LC, int LD>
QCG_HD void grad_weights(const PairRec& P, const PairRec& Q, const double* __restrict__ D, int nbf, double xfac,
                         double deg, double* W) {
    using S = EriShape<LA, LB, LC, LD>;
    constexpr bool U = S::kUnroll;
        // density weights of all components (norms folded in)
        for_n<U, S::NQ>([&](int k) {
            const int d = k % S::ND, c = (k / S::ND) % S::NC, b = (k / S::NCD) % S::NB, a = k / (S::NCD * S::NB);
            const int i = P.bfa + a, j = P.bfb + b, kk = Q.bfa + c, l = Q.bfb + d;
            const double dab = D[(size_t)i * nbf + j], dcd = D[(size_t)kk * nbf + l];
            const double dac = D[(size_t)i * nbf + kk], dbd = D[(size_t)j * nbf + l];
            const double dad = D[(size_t)i * nbf + l], dbc = D[(size_t)j * nbf + kk];
            W[k] = 4.0 * deg * (dab * dcd - xfac * (dac * dbd + dad * dbc)) * kCompNorm[LA][a] * kCompNorm[LB][b] *
                   kCompNorm[LC][c] * kCompNorm[LD][d];
        });
}

template <int LA, int LB, int LC, int LD>
QCG_HD void grad_quartet(const PairRec& P, const PairRec& Q, const PrimPair<double>* __restrict__ pp,
                         const double* __restrict__ rys_tab, const double* W, double* gq_out) {
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int EA = LA + 1, EB = LB + 1, EC = LC + 1, ED = LD;
    constexpr int NB_ = LA + LB + 2, NK_ = LC + LD + 1;
    constexpr int GE = (NB_ + 1) * (NK_ + 1);
    constexpr int HE = (EA + 1) * (EB + 1) * (EC + 1) * (ED + 1);
    constexpr int N = (NB_ + NK_) / 2 + 1;
    constexpr bool U = S::kUnroll;
    const double two_pi_52 = 2.0 * kPi * kPi * 1.7724538509055160273;
    auto hidx = [](int i, int j, int k, int l) { return ((i * (EB + 1) + j) * (EC + 1) + k) * (ED + 1) + l; };
        double gq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // A xyz, B xyz, C xyz
        for (int ip = 0; ip < P.npp; ++ip) {
            const PrimPair<double> a = pp[P.pp_off + ip];
            const double p = a.p, hp = 0.5 / p;
            const double al_a = a.pad0, al_b = a.pad1;
            for (int jq = 0; jq < Q.npp; ++jq) {
                const PrimPair<double> c = pp[Q.pp_off + jq];
                if (a.kab * c.kab < kPrimScreen) continue;
                const double qx = c.p, al_c = c.pad0;
                const double ps = p + qx, ipq = 1.0 / ps;
                const double PQx = a.px - c.px, PQy = a.py - c.py, PQz = a.pz - c.pz;
                const double T = p * qx * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
                const double pref = two_pi_52 / (p * qx * sqrt(ps)) * a.ck * c.ck;
                double rx[N], rw[N];
                rys_eval<double, N>(T, rys_tab, rx, rw);
                const double hq = 0.5 / qx, qf = qx * ipq, pf = p * ipq;
                const double PA[3] = {a.px - P.ax, a.py - P.ay, a.pz - P.az};
                const double QC[3] = {c.px - Q.ax, c.py - Q.ay, c.pz - Q.az};
                const double PQ[3] = {PQx, PQy, PQz};
                const double AB[3] = {P.abx, P.aby, P.abz}, CD[3] = {Q.abx, Q.aby, Q.abz};
#pragma unroll
                for (int r = 0; r < N; ++r) {
                    const double u = rx[r];
                    const double b00 = 0.5 * u * ipq, b10 = hp * (1.0 - u * qf), b01 = hq * (1.0 - u * pf);
                    double I[3][HE];
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        double g[GE];
                        vrr2d<NB_, NK_, double>(g, PA[x] - u * qf * PQ[x], QC[x] + u * pf * PQ[x], b00, b10, b01);
                        hrr1d<EA, EB, EC, ED, double>(g, AB[x], CD[x], I[x]);
                    }
                    const double wp = rw[r] * pref;
                    // derivative tables per axis and centre
                    double dA[3][S::H4], dB[3][S::H4], dC[3][S::H4], I0[3][S::H4];
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int i = 0; i <= LA; ++i)
#pragma unroll
                            for (int j = 0; j <= LB; ++j)
#pragma unroll
                                for (int k = 0; k <= LC; ++k)
#pragma unroll
                                    for (int l = 0; l <= LD; ++l) {
                                        const int h = ((i * (LB + 1) + j) * (LC + 1) + k) * (LD + 1) + l;
                                        I0[x][h] = I[x][hidx(i, j, k, l)];
                                        dA[x][h] = 2.0 * al_a * I[x][hidx(i + 1, j, k, l)] - (i ? i * I[x][hidx(i - 1, j, k, l)] : 0.0);
                                        dB[x][h] = 2.0 * al_b * I[x][hidx(i, j + 1, k, l)] - (j ? j * I[x][hidx(i, j - 1, k, l)] : 0.0);
                                        dC[x][h] = 2.0 * al_c * I[x][hidx(i, j, k + 1, l)] - (k ? k * I[x][hidx(i, j, k - 1, l)] : 0.0);
                                    }
                    constexpr const CompTab<LA, LB, LC, LD>& ct = kCompTab<LA, LB, LC, LD>;
                    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    for_n<U, S::NQ>([&](int k) {
                        const int hx = ct.hx[k], hy = ct.hy[k], hz = ct.hz[k];
                        const double w = W[k];
                        const double ix = I0[0][hx], iy = I0[1][hy], iz = I0[2][hz];
                        const double iyz = iy * iz, ixz = ix * iz, ixy = ix * iy;
                        acc[0] += w * dA[0][hx] * iyz;
                        acc[1] += w * dA[1][hy] * ixz;
                        acc[2] += w * dA[2][hz] * ixy;
                        acc[3] += w * dB[0][hx] * iyz;
                        acc[4] += w * dB[1][hy] * ixz;
                        acc[5] += w * dB[2][hz] * ixy;
                        acc[6] += w * dC[0][hx] * iyz;
                        acc[7] += w * dC[1][hy] * ixz;
                        acc[8] += w * dC[2][hz] * ixy;
                    });
#pragma unroll
                    for (int t = 0; t < 9; ++t) gq[t] += wp * acc[t];
                }
            }
        }
#pragma unroll
    for (int t = 0; t < 9; ++t) gq_out[t] = gq[t];
}

}  // namespace qcg
