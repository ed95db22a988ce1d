// Two-electron part of the analytic nuclear gradient (the adjoint of the Fock build) for all
// quartet classes whose bra pair class is QCG_CB (compiled once per bra class, like eri_jk.hip).
//
// Same quantity as qc-core's eri_gradient_fused (crates/qc-core/src/scf/gradient.rs:1430-1552) with
// compute_symmetry_weight (:1570-1627) and the derivative integrals of shell_eri_with_derivatives
// (integrals/eri/mod.rs:2035-2548):
//   dE/dR = sum over unique quartets, components: W_abcd d(ab|cd)/dR,
//   W_abcd = deg * 4 [ D_ab D_cd - c (D_ac D_bd + D_ad D_bc) ],  c = hf_frac / 4,
//   deg = 1/2 (a == b shell) * 1/2 (c == d shell) * 1/2 (pair P == pair Q)
// (the sum of 1/2 D_pq D_rs - c D_pr D_qs over the 8 index permutations).  Derivatives of a
// primitive with respect to its centre: d/dA_x phi_i = 2 alpha phi_(i+1) - i phi_(i-1), applied to
// the 1-D Rys factors for centres A, B, C; centre D by translational invariance, -(A + B + C).
//
// Wave-cooperative (eri_wave.h): one wavefront per work item (bra pair, range of kets), one
// contracted quartet at a time.  Phase A writes the 1-D transfer tables with powers up to
// (la+1, lb+1, lc+1, ld) for (L+1)/2 + 1 Rys nodes (the derivative integrals have total angular
// momentum L+1) and the three primitive exponents of every primitive quartet to LDS; phase B has
// the lanes over Cartesian components: W_k times the nine derivative products (A, B, C x xyz),
// summed over the entries.  No derivative integral is stored; the density weights W are formed
// once per quartet from density blocks staged in LDS.
#include "eri_launch.h"
#include "eri_wave.h"

#ifndef QCG_CB
#error "compile with -DQCG_CB=<bra pair class>"
#endif

namespace qcg {

template <int LA, int LB, int LC, int LD>
struct GradWave {
    using S = EriShape<LA, LB, LC, LD>;
    static constexpr int N = (LA + LB + LC + LD + 1) / 2 + 1;
    using T = WaveTable<LA + 1, LB + 1, LC + 1, LD, N, double>;
    static constexpr int DB = S::NAB + S::NCD + S::NA * S::NC + S::NB * S::ND + S::NA * S::ND + S::NB * S::NC;
    static constexpr int WAVE_D = T::DOUBLES + 3 * T::PQ_CAP + S::NQ + DB;
    // waves per block: four while their LDS fits in 64 KB, else two
    static constexpr int WPB = 4 * WAVE_D * 8 <= 65536 ? 4 : 2;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void grad_wave_kernel(const int4* __restrict__ items, const int2* __restrict__ segs, int nitems,
                                                        const PairRec* __restrict__ bra, const PairRec* __restrict__ ket,
                                                        int diag, const PrimPair<double>* __restrict__ pp,
                                                        const double* __restrict__ rys_tab,
                                                        const double* __restrict__ D, int nbf_, double xfac,
                                                        double* __restrict__ grad) {
    using S = EriShape<LA, LB, LC, LD>;
    using G = GradWave<LA, LB, LC, LD>;
    using TB = typename G::T;
    constexpr int N = G::N, HT = TB::HT, ES = TB::ES;
    constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND, NCD = S::NCD;
    constexpr int TB1 = LB + 2, TC1 = LC + 2, TD1 = LD + 1;
    constexpr int sI = TB1 * TC1 * TD1, sJ = TC1 * TD1, sK = TD1;
    __shared__ double smem[G::WPB * G::WAVE_D];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int item = blockIdx.x * G::WPB + wid;
    if (item >= nitems) return;  // waves are independent: no block-wide barrier below
    double* tbl = smem + wid * G::WAVE_D;
    double* alph = tbl + TB::DOUBLES;
    double* Wl = alph + 3 * TB::PQ_CAP;
    double* Dab = Wl + S::NQ;
    double* Dcd = Dab + S::NAB;
    double* Dac = Dcd + NCD;
    double* Dbd = Dac + NA * NC;
    double* Dad = Dbd + NB * ND;
    double* Dbc = Dad + NA * ND;
    const size_t nbf = (size_t)nbf_;

    const int4 it = items[item];
    const PairRec P = item_bra(bra, it);
    const int pidx = P.pos;
    KetCursor cur{segs, it.y, it.z};
    for (int o = lane; o < S::NAB; o += 64) Dab[o] = D[(P.bfa + o / NB) * nbf + P.bfb + o % NB];
    const double deg_p = (P.sa == P.sb) ? 0.5 : 1.0;
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};

    for (int j = 0; j < item_kets(it); ++j, ++cur.r) {
        const int q = cur.at();
        const PairRec Q = ket[q];
        const double deg = deg_p * ((Q.sa == Q.sb) ? 0.5 : 1.0) * ((diag && q == pidx) ? 0.5 : 1.0);
        wave_sync();  // previous quartet's readers of the weights and density blocks are done
        for (int o = lane; o < NCD; o += 64) Dcd[o] = D[(Q.bfa + o / ND) * nbf + Q.bfb + o % ND];
        for (int o = lane; o < NA * NC; o += 64) Dac[o] = D[(P.bfa + o / NC) * nbf + Q.bfa + o % NC];
        for (int o = lane; o < NB * ND; o += 64) Dbd[o] = D[(P.bfb + o / ND) * nbf + Q.bfb + o % ND];
        for (int o = lane; o < NA * ND; o += 64) Dad[o] = D[(P.bfa + o / ND) * nbf + Q.bfb + o % ND];
        for (int o = lane; o < NB * NC; o += 64) Dbc[o] = D[(P.bfb + o / NC) * nbf + Q.bfa + o % NC];
        wave_sync();
        for (int k = lane; k < S::NQ; k += 64) {
            const int d = k % ND, c = (k / ND) % NC, b = (k / NCD) % NB, a = k / (NCD * NB);
            Wl[k] = 4.0 * deg *
                    (Dab[a * NB + b] * Dcd[c * ND + d] - xfac * (Dac[a * NC + c] * Dbd[b * ND + d] + Dad[a * ND + d] * Dbc[b * NC + c])) *
                    cnorm(LA, a) * cnorm(LB, b) * cnorm(LC, c) * cnorm(LD, d);
        }
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // A xyz, B xyz, C xyz
        const int npq = P.npp * Q.npp;
        for (int pq0 = 0; pq0 < npq; pq0 += TB::PQ_CAP) {
            const int npc = min(TB::PQ_CAP, npq - pq0);
            wave_sync();
            wave_tables<LA + 1, LB + 1, LC + 1, LD, N, ES, double>(P, Q, pp, rys_tab, pq0, npc, tbl, lane);
            for (int u = lane; u < npc; u += 64) {
                const int pq = pq0 + u, i = pq / Q.npp, j = pq - i * Q.npp;
                const PrimPair<double>& a = pp[P.pp_off + i];
                alph[3 * u] = a.ea;
                alph[3 * u + 1] = a.p - a.ea;
                alph[3 * u + 2] = pp[Q.pp_off + j].ea;
            }
            wave_sync();
            const int ne = npc * N;
#pragma unroll 1
            for (int k = lane; k < S::NQ; k += 64) {
                const int d = k % ND, c = (k / ND) % NC, b = (k / NCD) % NB, a = k / (NCD * NB);
                const double w = Wl[k];
                // per axis: row pointers of I(i,j,k,l), I(i+1..), I(i-1..), I(.,j+1,..), ... and the powers
                const double* r0[3];
                const double *rAp[3], *rAm[3], *rBp[3], *rBm[3], *rCp[3], *rCm[3];
                double fi[3], fj[3], fk[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int ia = x == 0 ? cart_x(LA, a) : x == 1 ? cart_y(LA, a) : cart_z(LA, a);
                    const int ib = x == 0 ? cart_x(LB, b) : x == 1 ? cart_y(LB, b) : cart_z(LB, b);
                    const int ic = x == 0 ? cart_x(LC, c) : x == 1 ? cart_y(LC, c) : cart_z(LC, c);
                    const int id = x == 0 ? cart_x(LD, d) : x == 1 ? cart_y(LD, d) : cart_z(LD, d);
                    const int h0 = ((ia * TB1 + ib) * TC1 + ic) * TD1 + id;
                    const double* base = tbl + x * HT * ES;
                    r0[x] = base + h0 * ES;
                    rAp[x] = base + (h0 + sI) * ES;
                    rAm[x] = base + (ia ? h0 - sI : h0) * ES;
                    rBp[x] = base + (h0 + sJ) * ES;
                    rBm[x] = base + (ib ? h0 - sJ : h0) * ES;
                    rCp[x] = base + (h0 + sK) * ES;
                    rCm[x] = base + (ic ? h0 - sK : h0) * ES;
                    fi[x] = ia;
                    fj[x] = ib;
                    fk[x] = ic;
                }
                double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int e = 0; e < ne; ++e) {
                    const int pql = e / N;
                    const double two_a = 2.0 * alph[3 * pql], two_b = 2.0 * alph[3 * pql + 1],
                                 two_c = 2.0 * alph[3 * pql + 2];
                    const double i0x = r0[0][e], i0y = r0[1][e], i0z = r0[2][e];
                    const double iyz = i0y * i0z, ixz = i0x * i0z, ixy = i0x * i0y;
                    const double pr[3] = {iyz, ixz, ixy};
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        s[x] += (two_a * rAp[x][e] - fi[x] * rAm[x][e]) * pr[x];
                        s[3 + x] += (two_b * rBp[x][e] - fj[x] * rBm[x][e]) * pr[x];
                        s[6 + x] += (two_c * rCp[x][e] - fk[x] * rCm[x][e]) * pr[x];
                    }
                }
#pragma unroll
                for (int t = 0; t < 9; ++t) acc[t] += w * s[t];
            }
        }
        // ket atoms: C directly, D = -(A + B + C) of this quartet (translational invariance)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            gA[x] += acc[x];
            gB[x] += acc[3 + x];
            const double cx = wave_sum(acc[6 + x]);
            const double sx = wave_sum(acc[x] + acc[3 + x] + acc[6 + x]);
            if (lane == 0) {
                if (cx != 0.0) atomicAdd(&grad[3 * Q.atom_a + x], cx);
                if (sx != 0.0) atomicAdd(&grad[3 * Q.atom_b + x], -sx);
            }
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const double ax = wave_sum(gA[x]), bx = wave_sum(gB[x]);
        if (lane == 0) {
            if (ax != 0.0) atomicAdd(&grad[3 * P.atom_a + x], ax);
            if (bx != 0.0) atomicAdd(&grad[3 * P.atom_b + x], bx);
        }
    }
}

// ------------------------------------------------------------------ per-lane form (light classes)
// For classes with at most kGradLaneMaxNQ components one ket quartet per thread is cheaper than a
// wavefront per quartet (which would leave most lanes idle in phase B): octane/6-31+G* took 13.0
// instead of 26.0 ms per ERI gradient with 54 here instead of 27, even though the 54-component
// classes then use the whole register file (one wavefront per SIMD).  Every table index below is a
// compile-time constant (static_for), so the transfer tables stay in registers with no indexed
// register access.
constexpr int kGradLaneMaxNQ = 54;
constexpr int kGradLaneBlock = 128;

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void grad_quartet_lane(const PairRec& P, const PairRec& Q,
                                                  const PrimPair<double>* __restrict__ pp,
                                                  const double* __restrict__ rys_n, const double* W, double* gq) {
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int EA = LA + 1, EB = LB + 1, EC = LC + 1, ED = LD;
    constexpr int NB_ = EA + EB, NK_ = EC + ED;
    constexpr int GE = (NB_ + 1) * (NK_ + 1);
    constexpr int HE = (EA + 1) * (EB + 1) * (EC + 1) * (ED + 1);
    constexpr int N = (LA + LB + LC + LD + 1) / 2 + 1;
    constexpr int sI = (EB + 1) * (EC + 1) * (ED + 1), sJ = (EC + 1) * (ED + 1), sK = ED + 1;
    const double two_pi_52 = 2.0 * kPi * kPi * 1.7724538509055160273;
    const double AB[3] = {P.abx, P.aby, P.abz}, CD[3] = {Q.abx, Q.aby, Q.abz};
    for (int j = 0; j < Q.npp; ++j) {
        const PrimPair<double> c = pp[Q.pp_off + j];
        const double q = c.p, hq = c.hp, al_c = c.ea;
        const double QC[3] = {c.px - Q.ax, c.py - Q.ay, c.pz - Q.az};
        for (int i = 0; i < P.npp; ++i) {
            const PrimPair<double> a = pp[P.pp_off + i];
            if (a.kab * c.kab < kPrimScreen) continue;
            const double p = a.p, hp = a.hp, al_a = a.ea, al_b = a.p - a.ea;
            const double rs = qrsqrt(p + q), ipq = rs * rs;
            const double PQ[3] = {a.px - c.px, a.py - c.py, a.pz - c.pz};
            const double T = p * q * ipq * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = two_pi_52 * rs * a.ckp * c.ckp;
            const double PA[3] = {a.px - P.ax, a.py - P.ay, a.pz - P.az};
            double rx[N], rw[N];
            rys_eval_tab<double, N>(T, rys_n, rx, rw);
            const double qf = q * ipq, pf = p * ipq;
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
                double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                static_for<S::NQ>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    constexpr int d = k % S::ND, cc = (k / S::ND) % S::NC, b = (k / S::NCD) % S::NB, aa = k / (S::NCD * S::NB);
                    constexpr int ia[3] = {cart_x(LA, aa), cart_y(LA, aa), cart_z(LA, aa)};
                    constexpr int ib[3] = {cart_x(LB, b), cart_y(LB, b), cart_z(LB, b)};
                    constexpr int ic[3] = {cart_x(LC, cc), cart_y(LC, cc), cart_z(LC, cc)};
                    constexpr int id[3] = {cart_x(LD, d), cart_y(LD, d), cart_z(LD, d)};
                    constexpr int h0[3] = {((ia[0] * (EB + 1) + ib[0]) * (EC + 1) + ic[0]) * (ED + 1) + id[0],
                                           ((ia[1] * (EB + 1) + ib[1]) * (EC + 1) + ic[1]) * (ED + 1) + id[1],
                                           ((ia[2] * (EB + 1) + ib[2]) * (EC + 1) + ic[2]) * (ED + 1) + id[2]};
                    const double w = W[k];
                    const double i0[3] = {I[0][h0[0]], I[1][h0[1]], I[2][h0[2]]};
                    const double pr[3] = {i0[1] * i0[2], i0[0] * i0[2], i0[0] * i0[1]};
                    static_for<3>([&](auto xc) {
                        constexpr int x = decltype(xc)::value;
                        double dA = 2.0 * al_a * I[x][h0[x] + sI];
                        if constexpr (ia[x] > 0) dA -= ia[x] * I[x][h0[x] - sI];
                        double dB = 2.0 * al_b * I[x][h0[x] + sJ];
                        if constexpr (ib[x] > 0) dB -= ib[x] * I[x][h0[x] - sJ];
                        double dC = 2.0 * al_c * I[x][h0[x] + sK];
                        if constexpr (ic[x] > 0) dC -= ic[x] * I[x][h0[x] - sK];
                        acc[x] += w * dA * pr[x];
                        acc[3 + x] += w * dB * pr[x];
                        acc[6 + x] += w * dC * pr[x];
                    });
                });
                const double wp = rw[r] * pref;
#pragma unroll
                for (int t = 0; t < 9; ++t) gq[t] += wp * acc[t];
            }
        }
    }
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(kGradLaneBlock) void grad_lane_kernel(
    const int4* __restrict__ items, const int2* __restrict__ segs, const PairRec* __restrict__ bra, const PairRec* __restrict__ ket, int diag,
    const PrimPair<double>* __restrict__ pp, const double* __restrict__ rys_tab, const double* __restrict__ D,
    int nbf_, double xfac, int natom, double* __restrict__ grad) {
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int N = (LA + LB + LC + LD + 1) / 2 + 1;
    constexpr int RT = rys_tab_size<N>();
    __shared__ double rys_s[RT];
    extern __shared__ double gl[];  // natom x 3: ket-atom sums of the block
    const int tid = threadIdx.x;
    for (int i = tid; i < RT; i += blockDim.x) rys_s[i] = rys_tab[qcg_rys_off[N] + i];
    for (int i = tid; i < 3 * natom; i += blockDim.x) gl[i] = 0.0;
    __syncthreads();
    const size_t nbf = (size_t)nbf_;
    const int4 it = items[blockIdx.x];
    const PairRec P = item_bra(bra, it);
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    if (tid < item_kets(it)) {
        KetCursor cur{segs, it.y, it.z + tid};
        const int q = cur.at();
        const PairRec Q = ket[q];
        const double deg = ((P.sa == P.sb) ? 0.5 : 1.0) * ((Q.sa == Q.sb) ? 0.5 : 1.0) * ((diag && q == P.pos) ? 0.5 : 1.0);
        double W[S::NQ];
        static_for<S::NQ>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int d = k % S::ND, c = (k / S::ND) % S::NC, b = (k / S::NCD) % S::NB, a = k / (S::NCD * S::NB);
            const size_t i_ = P.bfa + a, j_ = P.bfb + b, k_ = Q.bfa + c, l_ = Q.bfb + d;
            W[k] = 4.0 * deg *
                   (D[i_ * nbf + j_] * D[k_ * nbf + l_] - xfac * (D[i_ * nbf + k_] * D[j_ * nbf + l_] + D[i_ * nbf + l_] * D[j_ * nbf + k_])) *
                   (cnorm(LA, a) * cnorm(LB, b) * cnorm(LC, c) * cnorm(LD, d));
        });
        double gq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        grad_quartet_lane<LA, LB, LC, LD>(P, Q, pp, rys_s, W, gq);
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            gA[x] = gq[x];
            gB[x] = gq[3 + x];
            atomicAdd(&gl[3 * Q.atom_a + x], gq[6 + x]);
            atomicAdd(&gl[3 * Q.atom_b + x], -(gq[x] + gq[3 + x] + gq[6 + x]));  // centre D
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const double ax = wave_sum(gA[x]), bx = wave_sum(gB[x]);
        if ((tid & 63) == 0) {
            atomicAdd(&gl[3 * P.atom_a + x], ax);
            atomicAdd(&gl[3 * P.atom_b + x], bx);
        }
    }
    __syncthreads();
    for (int i = tid; i < 3 * natom; i += blockDim.x)
        if (gl[i] != 0.0) atomicAdd(&grad[i], gl[i]);
}

#define QCG_LA(c) ((c) == 0 ? 0 : (c) <= 2 ? 1 : 2)
#define QCG_LB(c) ((c) == 0 || (c) == 1 || (c) == 3 ? 0 : ((c) == 2 || (c) == 4) ? 1 : 2)

template <int CB, int CK>
static hipError_t launch_grad_t(const GradLaunch& a, hipStream_t st) {
    constexpr int LA = QCG_LA(CB), LB = QCG_LB(CB), LC = QCG_LA(CK), LD = QCG_LB(CK);
    using G = GradWave<LA, LB, LC, LD>;
    if (a.nitems <= 0) return hipSuccess;
    if constexpr (EriShape<LA, LB, LC, LD>::NQ <= kGradLaneMaxNQ) {
        hipLaunchKernelGGL((grad_lane_kernel<LA, LB, LC, LD>), dim3(a.nitems), dim3(kGradLaneBlock),
                           3 * (size_t)a.natom * sizeof(double), st, a.items, a.segs, a.bra, a.ket, a.diag, a.pp, a.rys, a.D,
                           a.nbf, a.xfac, a.natom, a.grad);
    } else {
        hipLaunchKernelGGL((grad_wave_kernel<LA, LB, LC, LD>), dim3((a.nitems + G::WPB - 1) / G::WPB),
                           dim3(64 * G::WPB), 0, st, a.items, a.segs, a.nitems, a.bra, a.ket, a.diag, a.pp, a.rys, a.D, a.nbf,
                           a.xfac, a.grad);
    }
    return hipGetLastError();
}

#define QCG_CAT2(a, b) a##b
#define QCG_CAT(a, b) QCG_CAT2(a, b)

hipError_t QCG_CAT(launch_grad_cb, QCG_CB)(int ck, const GradLaunch& a, hipStream_t st) {
    switch (ck) {
#define QCG_K(CK) \
    case CK: if constexpr (CK <= QCG_CB) return launch_grad_t<QCG_CB, CK>(a, st); break;
        QCG_K(0) QCG_K(1) QCG_K(2) QCG_K(3) QCG_K(4) QCG_K(5)
#undef QCG_K
    }
    return hipErrorInvalidValue;
}

}  // namespace qcg
