// Direct J/K build and compressed-tensor kernels for all quartet classes whose bra pair class
// is QCG_CB (this file is compiled once per bra class, -DQCG_CB=0..5, so the six units build
// in parallel).  Pair class c <-> (la, lb): 0 (s,s) 1 (p,s) 2 (p,p) 3 (d,s) 4 (d,p) 5 (d,d).
//
// Work decomposition of the J/K build (one launch per quartet class, bra class >= ket class):
//   * pairs of a class are ordered by contraction bucket (number of primitive pairs), then by
//     Schwarz bound (descending); for a bra pair P the surviving kets of a bucket are the prefix
//     with Q_P*Q_Q >= threshold (monotone product), found by binary search on exactly the product
//     the reference tests (q_ij * q_kl < SCHWARZ_THRESHOLD, integrals/eri/mod.rs:1008-1023); for
//     bra class == ket class only kets at or before P (each unordered pair-pair once = 8-fold
//     symmetry); in mixed precision each prefix splits into an FP64 range (Q_P*Q_Q >= cutoff) and
//     an FP32 range;
//   * the host cuts every range into work items (bra pair, ket sub-range) of a per-class size;
//     one block per item, one thread per ket (several for cheap classes), so every class
//     launch has many more blocks than the GPU has CUs;
//   * digestion: J_P in registers + wave reduction; J_Q by FP64 atomics; the four K terms all
//     land in the rows of the bra functions, so they are summed in LDS rows (one block = one bra
//     pair) and flushed with one global atomic per touched element at the end of the block (device
//     atomics are performed memory-side on this GPU, so every global atomic is HBM write
//     traffic).  J = Jacc + Jacc^T, K = Kacc + Kacc^T after.
//   * symmetry factors: 1/2 for a == b shells, 1/2 for c == d shells, 1/2 for P == Q, folded
//     together with the component normalisation into the density sub-blocks.
// With the symmetric density this reproduces build_fock's G = J - K/2 (scf/mod.rs:825-881) and
// build_jk (dft/ks_scf.rs:1369-1394).
#include "eri_device.h"
#include "eri_launch.h"
#include "eri_wave.h"

#ifndef QCG_CB
#error "compile with -DQCG_CB=<bra pair class>"
#endif
// Per-lane kernel options, both off by default: the bra's primitive pairs staged in LDS (off: the
// block-uniform reads go through the scalar cache), the D rows of the bra functions staged in LDS
// next to the K rows (off: the K digestion reads D through the caches).  The smaller LDS footprint
// keeps more blocks resident: decane, with tiles retuned for each, 643 -> 652 builds/s FP64 and
// 631 -> 643 mixed without the D rows, then 652 -> 677 and 643 -> 659 without the bra staging.
#ifndef QCG_JK_BRA_LDS
#define QCG_JK_BRA_LDS 0
#endif
#ifndef QCG_JK_DROWS
#define QCG_JK_DROWS 0
#endif

namespace qcg {

// The Rys coefficients of order N for a block: staged in LDS (lds, at least rys_tab_size<N>()
// elements) when N <= 3, where the per-lane kernels gather them at lane-dependent intervals once
// per primitive quartet, else read from device memory.  The caller's __syncthreads() follows.
template <typename R, int N>
__device__ __forceinline__ const R* rys_stage(const R* __restrict__ rys_tab, R* lds) {
    if constexpr (N <= 3) {
        for (int i = threadIdx.x; i < rys_tab_size<N>(); i += blockDim.x) lds[i] = rys_tab[qcg_rys_off[N] + i];
        return lds;
    } else {
        return rys_tab + qcg_rys_off[N];
    }
}
template <int N>
constexpr int rys_lds_doubles() { return N <= 3 ? rys_tab_size<N>() : 1; }

// One work item = (bra pair, run of its surviving kets over the ket segments, KetCursor); the
// host builds the items with the exact screening predicate, so no quartet is tested twice or
// skipped.  rys_n: the coefficients of this class's Rys order (rys_stage); bp: LDS room for the
// bra's primitive pairs; rows: LDS room for the K rows of the bra functions a and b (krow bit 0:
// (NA + NB) x nbf fixed-point words, fixacc.h, in which the four K terms of the item, all in the
// rows of a and b, are summed and then flushed once per block) followed by their D rows (bit 1:
// (NA + NB) x nbf doubles, read by the K digestion).  J and K go to the fixed-point accumulators
// Jacc/Kacc with the scales fix[4 mol], fix[4 mol + 1], so the result does not depend on the order
// in which blocks and lanes arrive.
template <int LA, int LB, int LC, int LD, typename R, int ROWS>
__device__ __forceinline__ void jk_item(
    const int4 it, const int2* __restrict__ segs, const PairRec* __restrict__ bra, const PairRec* __restrict__ ket, int diag,
    const PrimPair<R>* __restrict__ pp, const R* __restrict__ rys_n, const double* __restrict__ D,
    fix_t* __restrict__ Jacc, fix_t* __restrict__ Kacc, const double* __restrict__ fix, int do_j, int do_k,
    const int* __restrict__ ket_off, const int64_t* __restrict__ mat_off, const int* __restrict__ mol_nbf,
    PrimPair<R>* bp, double* rows) {
#if !QCG_JK_BRA_LDS
    (void)bp;
#endif
    using S = EriShape<LA, LB, LC, LD>;
    constexpr bool U = S::kUnroll;
    constexpr int NR = S::NA + S::NB;
    const PairRec P = item_bra(bra, it);
    const int pidx = P.pos;
    ket += ket_off[P.mol];
    D += mat_off[P.mol];
    Jacc += mat_off[P.mol];
    Kacc += mat_off[P.mol];
    const int nbf = mol_nbf[P.mol];
    const double sJ = fix[4 * P.mol], sK = fix[4 * P.mol + 1];  // fixed-point scales of J and K
    const int tid = threadIdx.x;
#if QCG_JK_BRA_LDS
    for (int i = tid; i < P.npp; i += blockDim.x) bp[i] = pp[P.pp_off + i];
    const PrimPair<R>* ppb = bp;
#else
    const PrimPair<R>* ppb = pp + P.pp_off;
#endif
    // ROWS bit 0: K rows in LDS, bit 1: D rows in LDS after them (compile-time, so each
    // instantiation holds one addressing path only)
    constexpr bool kr = ROWS & 1, dr = ROWS & 2;
    const int nrow = NR * nbf;
    fix_t* kl = reinterpret_cast<fix_t*>(rows);
    double* dl = rows + (kr ? nrow : 0);
    if constexpr (ROWS != 0)
        for (int i = tid; i < nrow; i += blockDim.x) {
            if constexpr (dr) {
                const int r = i / nbf, col = i - r * nbf;
                dl[i] = D[(size_t)(r < S::NA ? P.bfa + r : P.bfb + (r - S::NA)) * nbf + col];
            }
            if constexpr (kr) kl[i] = 0;
        }
    __syncthreads();
    // D element (row function a + r of the bra's shell a, or b + r of shell b; column col)
    auto drow = [&](bool row_b, int r, int col) -> double {
        if constexpr (dr) return dl[(row_b ? S::NA + r : r) * nbf + col];
        else return D[(size_t)((row_b ? P.bfb : P.bfa) + r) * nbf + col];
    };
    auto kadd = [&](bool row_b, int r, int col, double v) {
        if constexpr (kr) fix_add_lds(kl, (row_b ? S::NA + r : r) * nbf + col, v, sK);
        else fix_add_global(Kacc, (size_t)((row_b ? P.bfb : P.bfa) + r) * nbf + col, v, sK);
    };

    // bra density block scaled by the bra component norms
    double Dps[S::NAB], jp[S::NAB];
    for_n<U, S::NAB>([&](int ab) {
        const int a = ab / S::NB, b = ab % S::NB;
        Dps[ab] = drow(false, a, P.bfb + b) * cnorm(LA, a) * cnorm(LB, b);
        jp[ab] = 0.0;
    });
    const double deg_p = (P.sa == P.sb) ? 0.5 : 1.0;

    KetCursor cur{segs, it.y, it.z + tid};
    for (int j = tid; j < item_kets(it); j += blockDim.x, cur.r += blockDim.x) {
        const int q = cur.at();
        const PairRec Q = ket[q];
        R eri[S::NQ];
        for_n<U, S::NQ>([&](int i) { eri[i] = R(0); });
        contract_quartet_pp<LA, LB, LC, LD, R, true>(P, Q, ppb, pp + Q.pp_off, rys_n, eri);
        const double deg = deg_p * ((Q.sa == Q.sb) ? 0.5 : 1.0) * ((diag && q == pidx) ? 0.5 : 1.0);
        if (do_j) {
            double Dqs[S::NCD];
            for_n<U, S::NCD>([&](int cd) {
                const int c = cd / S::ND, d = cd % S::ND;
                Dqs[cd] = D[(size_t)(Q.bfa + c) * nbf + Q.bfb + d] * cnorm(LC, c) * cnorm(LD, d);
            });
            const double fj = 2.0 * deg;
            for_n<U, S::NAB>([&](int ab) {
                double s = 0.0;
                for_n<U, S::NCD>([&](int cd) { s += Dqs[cd] * (double)eri[ab * S::NCD + cd]; });
                const int a = ab / S::NB, b = ab % S::NB;
                jp[ab] += fj * cnorm(LA, a) * cnorm(LB, b) * s;
            });
            for_n<U, S::NCD>([&](int cd) {
                double s = 0.0;
                for_n<U, S::NAB>([&](int ab) { s += Dps[ab] * (double)eri[ab * S::NCD + cd]; });
                const int c = cd / S::ND, d = cd % S::ND;
                fix_add_global(Jacc, (size_t)(Q.bfa + c) * nbf + Q.bfb + d, fj * cnorm(LC, c) * cnorm(LD, d) * s, sJ);
            });
        }
        if (do_k) {
            // Kacc_ik += D_jl v, Kacc_jk += D_il v, Kacc_il += D_jk v, Kacc_jl += D_ik v
            // (i in a, j in b, k in c, l in d); every D sub-block pre-scaled by the norms it sums over.
            double Dbd[S::NB * S::ND], Dad[S::NA * S::ND], Dbc[S::NB * S::NC], Dac[S::NA * S::NC];
            for_n<U, S::NB * S::ND>([&](int i) {
                const int b = i / S::ND, d = i % S::ND;
                Dbd[i] = drow(true, b, Q.bfb + d) * cnorm(LB, b) * cnorm(LD, d);
            });
            for_n<U, S::NA * S::ND>([&](int i) {
                const int a = i / S::ND, d = i % S::ND;
                Dad[i] = drow(false, a, Q.bfb + d) * cnorm(LA, a) * cnorm(LD, d);
            });
            for_n<U, S::NB * S::NC>([&](int i) {
                const int b = i / S::NC, c = i % S::NC;
                Dbc[i] = drow(true, b, Q.bfa + c) * cnorm(LB, b) * cnorm(LC, c);
            });
            for_n<U, S::NA * S::NC>([&](int i) {
                const int a = i / S::NC, c = i % S::NC;
                Dac[i] = drow(false, a, Q.bfa + c) * cnorm(LA, a) * cnorm(LC, c);
            });
#define VIDX(a, b, c, d) ((((a) * S::NB + (b)) * S::NC + (c)) * S::ND + (d))
            for_n<U, S::NA * S::NC>([&](int ac) {
                const int a = ac / S::NC, c = ac % S::NC;
                double s = 0.0;
                for_n<U, S::NB * S::ND>([&](int bd) {
                    const int b = bd / S::ND, d = bd % S::ND;
                    s += Dbd[bd] * (double)eri[VIDX(a, b, c, d)];
                });
                kadd(false, a, Q.bfa + c, deg * cnorm(LA, a) * cnorm(LC, c) * s);
            });
            for_n<U, S::NB * S::NC>([&](int bc) {
                const int b = bc / S::NC, c = bc % S::NC;
                double s = 0.0;
                for_n<U, S::NA * S::ND>([&](int ad) {
                    const int a = ad / S::ND, d = ad % S::ND;
                    s += Dad[ad] * (double)eri[VIDX(a, b, c, d)];
                });
                kadd(true, b, Q.bfa + c, deg * cnorm(LB, b) * cnorm(LC, c) * s);
            });
            for_n<U, S::NA * S::ND>([&](int ad) {
                const int a = ad / S::ND, d = ad % S::ND;
                double s = 0.0;
                for_n<U, S::NB * S::NC>([&](int bc) {
                    const int b = bc / S::NC, c = bc % S::NC;
                    s += Dbc[bc] * (double)eri[VIDX(a, b, c, d)];
                });
                kadd(false, a, Q.bfb + d, deg * cnorm(LA, a) * cnorm(LD, d) * s);
            });
            for_n<U, S::NB * S::ND>([&](int bd) {
                const int b = bd / S::ND, d = bd % S::ND;
                double s = 0.0;
                for_n<U, S::NA * S::NC>([&](int ac) {
                    const int a = ac / S::NC, c = ac % S::NC;
                    s += Dac[ac] * (double)eri[VIDX(a, b, c, d)];
                });
                kadd(true, b, Q.bfb + d, deg * cnorm(LB, b) * cnorm(LD, d) * s);
            });
#undef VIDX
        }
    }
    // J_P: wave reduction (a fixed shuffle tree), one fixed-point add per component per wave
    if (do_j) {
        for_n<true, S::NAB>([&](int ab) {
            double s = jp[ab];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
            if ((tid & 63) == 0 && s != 0.0) {
                const int a = ab / S::NB, b = ab % S::NB;
                fix_add_global(Jacc, (size_t)(P.bfa + a) * nbf + P.bfb + b, s, sJ);
            }
        });
    }
    if constexpr (kr) {
        __syncthreads();
        for (int i = tid; i < nrow; i += blockDim.x) {
            const int r = i / nbf, col = i - r * nbf;
            const int row = r < S::NA ? P.bfa + r : P.bfb + (r - S::NA);
            fix_add_word_global(Kacc, (size_t)row * nbf + col, kl[i]);
        }
    }
}

// LDS layout of the per-lane kernels' rows for nbf functions (the largest molecule's): K rows
// as fixed-point words, (NA + NB) x nbf x 8 bytes, when they fit the budget left by the static
// LDS (Rys coefficients, bra primitives), then the D rows, (NA + NB) x nbf x 8 bytes, when those
// fit too.  Returns the krow flags (bit 0 K rows, bit 1 D rows) and the dynamic LDS bytes.
constexpr size_t kLdsBudget = 65536;
template <int LA, int LB, int LC, int LD>
inline int jk_row_flags(int nbf, int do_k, int bra_pp, size_t* bytes) {
    using S = EriShape<LA, LB, LC, LD>;
    const size_t fixed = (size_t)rys_lds_doubles<S::NROOTS>() * sizeof(double) + (size_t)bra_pp * sizeof(PrimPair<double>);
    const size_t kb = (size_t)(S::NA + S::NB) * nbf * sizeof(fix_t), db = (size_t)(S::NA + S::NB) * nbf * sizeof(double);
    int f = 0;
    size_t n = 0;
    if (do_k && fixed + kb <= kLdsBudget) {
        f |= 1;
        n += kb;
    }
    if (QCG_JK_DROWS && (f & 1) && fixed + n + db <= kLdsBudget) {
        f |= 2;
        n += db;
    }
    *bytes = n + (size_t)bra_pp * sizeof(PrimPair<double>);
    return f;
}

template <int LA, int LB, int LC, int LD, typename R, int ROWS>
__global__ __launch_bounds__(256) void jk_kernel(
    const int4* __restrict__ items, const int2* __restrict__ segs, const PairRec* __restrict__ bra, const PairRec* __restrict__ ket, int diag,
    const PrimPair<R>* __restrict__ pp, const R* __restrict__ rys_tab, const double* __restrict__ D,
    fix_t* __restrict__ Jacc, fix_t* __restrict__ Kacc, const double* __restrict__ fix, int do_j, int do_k,
    const int* __restrict__ ket_off, const int64_t* __restrict__ mat_off, const int* __restrict__ mol_nbf,
    int bra_pp) {
    constexpr int NR = EriShape<LA, LB, LC, LD>::NROOTS;
    extern __shared__ double dyn[];  // bra primitive pairs (bra_pp), then the K / D rows
    __shared__ double rys_raw[rys_lds_doubles<NR>()];
    const R* rys_n = rys_stage<R, NR>(rys_tab, reinterpret_cast<R*>(rys_raw));
    jk_item<LA, LB, LC, LD, R, ROWS>(items[blockIdx.x], segs, bra, ket, diag, pp, rys_n, D, Jacc, Kacc, fix, do_j, do_k,
                                     ket_off, mat_off, mol_nbf, reinterpret_cast<PrimPair<R>*>(dyn),
                               dyn + bra_pp * sizeof(PrimPair<R>) / sizeof(double));
}

// Mixed precision in one launch: items [0, n64) are FP64 quartets, the rest FP32 (one branch per
// block); saves the second launch per class and its tail.
template <int LA, int LB, int LC, int LD, int ROWS>
__global__ __launch_bounds__(256) void jk_kernel_mp(
    const int4* __restrict__ items, const int2* __restrict__ segs, int n64, const PairRec* __restrict__ bra, const PairRec* __restrict__ ket,
    int diag, const PrimPair<double>* __restrict__ pp, const PrimPair<float>* __restrict__ ppf,
    const double* __restrict__ rys_tab, const float* __restrict__ rys_tabf, const double* __restrict__ D,
    fix_t* __restrict__ Jacc, fix_t* __restrict__ Kacc, const double* __restrict__ fix, int do_j, int do_k,
    const int* __restrict__ ket_off, const int64_t* __restrict__ mat_off, const int* __restrict__ mol_nbf,
    int bra_pp) {
    constexpr int NR = EriShape<LA, LB, LC, LD>::NROOTS;
    extern __shared__ double dyn[];  // bra primitive pairs (bra_pp FP64 slots), then the K / D rows
    __shared__ double rys_raw[rys_lds_doubles<NR>()];
    double* rows = dyn + bra_pp * sizeof(PrimPair<double>) / sizeof(double);
    if ((int)blockIdx.x < n64) {
        const double* rys_n = rys_stage<double, NR>(rys_tab, rys_raw);
        jk_item<LA, LB, LC, LD, double, ROWS>(items[blockIdx.x], segs, bra, ket, diag, pp, rys_n, D, Jacc, Kacc, fix,
                                              do_j, do_k, ket_off, mat_off, mol_nbf,
                                              reinterpret_cast<PrimPair<double>*>(dyn), rows);
    } else {
        const float* rys_n = rys_stage<float, NR>(rys_tabf, reinterpret_cast<float*>(rys_raw));
        jk_item<LA, LB, LC, LD, float, ROWS>(items[blockIdx.x], segs, bra, ket, diag, ppf, rys_n, D, Jacc, Kacc, fix,
                                             do_j, do_k, ket_off, mat_off, mol_nbf,
                                             reinterpret_cast<PrimPair<float>*>(dyn), rows);
    }
}

// Wave-cooperative J/K for the classes with many components (eri_wave.h): four wavefronts per
// block, each takes one work item (bra pair, ket range) and walks its kets one quartet at a time.
// Digestion of a contracted quartet (phase C): the quartet v (norms and symmetry factor applied)
// goes to LDS, lanes over output elements form
//   J_ab += 2 sum_cd D_cd v   (kept in a register of lane ab until the item ends),
//   J_cd += 2 sum_ab D_ab v,  K_ac += sum_bd D_bd v,  K_bc += sum_ad D_ad v,
//   K_ad += sum_bc D_bc v,    K_bd += sum_ac D_ac v   (FP64 atomics per quartet).
template <int LA, int LB, int LC, int LD, typename R>
struct JkWave {
    using S = EriShape<LA, LB, LC, LD>;
    using T = WaveTable<LA, LB, LC, LD, S::NROOTS, R>;
    static constexpr int VD = T::DOUBLES > S::NQ ? T::DOUBLES : S::NQ;  // table, then the quartet
    static constexpr int DB = S::NAB + S::NCD + S::NB * S::ND + S::NA * S::ND + S::NB * S::NC + S::NA * S::NC;
    static constexpr int WAVE_D = VD + DB;
    static constexpr int KPL = (S::NQ + 63) / 64;
    static constexpr bool kEnabled = S::NQ >= 27;
};

// One work item on one wavefront; V: the wave's LDS region (JkWave<..., R>::WAVE_D doubles).
template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void jk_wave_item(
    const int4 it, const int2* __restrict__ segs, int lane, double* V, const PairRec* __restrict__ bra, const PairRec* __restrict__ ket,
    int diag, const PrimPair<R>* __restrict__ pp, const R* __restrict__ rys_tab, const double* __restrict__ D,
    fix_t* __restrict__ Jacc, fix_t* __restrict__ Kacc, const double* __restrict__ fix, int do_j, int do_k,
    const int* __restrict__ ket_off, const int64_t* __restrict__ mat_off, const int* __restrict__ mol_nbf) {
    using S = EriShape<LA, LB, LC, LD>;
    using W = JkWave<LA, LB, LC, LD, R>;
    using TB = typename W::T;
    constexpr int N = S::NROOTS, HT = TB::HT, ES = TB::ES, KPL = W::KPL;
    constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND, NCD = S::NCD;
    R* tbl = reinterpret_cast<R*>(V);
    double* Dab = V + W::VD;
    double* Dcd = Dab + S::NAB;
    double* Dbd = Dcd + NCD;
    double* Dad = Dbd + NB * ND;
    double* Dbc = Dad + NA * ND;
    double* Dac = Dbc + NB * NC;

    const PairRec P = item_bra(bra, it);
    const int pidx = P.pos;
    ket += ket_off[P.mol];
    D += mat_off[P.mol];
    Jacc += mat_off[P.mol];
    Kacc += mat_off[P.mol];
    const size_t nbf = (size_t)mol_nbf[P.mol];
    const double sJ = fix[4 * P.mol], sK = fix[4 * P.mol + 1];  // fixed-point scales of J and K

    constexpr const CompTab<LA, LB, LC, LD>& ct = kCompTab<LA, LB, LC, LD>;
    int ox[KPL], oy[KPL], oz[KPL];
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        const int k = lane + 64 * j;
        const int kk = k < S::NQ ? k : 0;
        ox[j] = ct.hx[kk] * ES;
        oy[j] = (HT + ct.hy[kk]) * ES;
        oz[j] = (2 * HT + ct.hz[kk]) * ES;
    }
    for (int o = lane; o < S::NAB; o += 64) Dab[o] = D[(P.bfa + o / NB) * nbf + P.bfb + o % NB];
    double jp = 0.0;  // J_ab of lane ab over the item's kets
    const double deg_p = (P.sa == P.sb) ? 0.5 : 1.0;

    KetCursor cur{segs, it.y, it.z};
    for (int j = 0; j < item_kets(it); ++j, ++cur.r) {
        const int q = cur.at();
        const PairRec Q = ket[q];
        R acc[KPL];
#pragma unroll
        for (int j = 0; j < KPL; ++j) acc[j] = R(0);
        const int npq = P.npp * Q.npp;
        for (int pq0 = 0; pq0 < npq; pq0 += TB::PQ_CAP) {
            const int npc = min(TB::PQ_CAP, npq - pq0);
            wave_sync();  // earlier readers of the table (and of V, which aliases it) are done
            wave_tables<LA, LB, LC, LD, N, ES, R>(P, Q, pp, rys_tab, pq0, npc, tbl, lane);
            wave_sync();
            const int ne = npc * N;
            for (int e = 0; e < ne; ++e) {
#pragma unroll
                for (int j = 0; j < KPL; ++j) acc[j] += tbl[ox[j] + e] * tbl[oy[j] + e] * tbl[oz[j] + e];
            }
        }
        wave_sync();
        const double deg = deg_p * ((Q.sa == Q.sb) ? 0.5 : 1.0) * ((diag && q == pidx) ? 0.5 : 1.0);
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            const int k = lane + 64 * j;
            if (k < S::NQ) {
                const int d = k % ND, c = (k / ND) % NC, b = (k / NCD) % NB, a = k / (NCD * NB);
                V[k] = (double)acc[j] * (cnorm(LA, a) * cnorm(LB, b) * cnorm(LC, c) * cnorm(LD, d) * deg);
            }
        }
        if (do_j)
            for (int o = lane; o < NCD; o += 64) Dcd[o] = D[(Q.bfa + o / ND) * nbf + Q.bfb + o % ND];
        if (do_k) {
            for (int o = lane; o < NB * ND; o += 64) Dbd[o] = D[(P.bfb + o / ND) * nbf + Q.bfb + o % ND];
            for (int o = lane; o < NA * ND; o += 64) Dad[o] = D[(P.bfa + o / ND) * nbf + Q.bfb + o % ND];
            for (int o = lane; o < NB * NC; o += 64) Dbc[o] = D[(P.bfb + o / NC) * nbf + Q.bfa + o % NC];
            for (int o = lane; o < NA * NC; o += 64) Dac[o] = D[(P.bfa + o / NC) * nbf + Q.bfa + o % NC];
        }
        wave_sync();
#define VAT(a, b, c, d) V[(((a) * NB + (b)) * NC + (c)) * ND + (d)]
        if (do_j) {
            if (lane < S::NAB) {
                double s = 0.0;
                for (int cd = 0; cd < NCD; ++cd) s += Dcd[cd] * V[lane * NCD + cd];
                jp += 2.0 * s;
            }
            for (int o = lane; o < NCD; o += 64) {
                double s = 0.0;
                for (int ab = 0; ab < S::NAB; ++ab) s += Dab[ab] * V[ab * NCD + o];
                fix_add_global(Jacc, (Q.bfa + o / ND) * nbf + Q.bfb + o % ND, 2.0 * s, sJ);
            }
        }
        if (do_k) {
            for (int o = lane; o < NA * NC; o += 64) {
                const int a = o / NC, c = o % NC;
                double s = 0.0;
                for (int b = 0; b < NB; ++b)
                    for (int d = 0; d < ND; ++d) s += Dbd[b * ND + d] * VAT(a, b, c, d);
                fix_add_global(Kacc, (P.bfa + a) * nbf + Q.bfa + c, s, sK);
            }
            for (int o = lane; o < NB * NC; o += 64) {
                const int b = o / NC, c = o % NC;
                double s = 0.0;
                for (int a = 0; a < NA; ++a)
                    for (int d = 0; d < ND; ++d) s += Dad[a * ND + d] * VAT(a, b, c, d);
                fix_add_global(Kacc, (P.bfb + b) * nbf + Q.bfa + c, s, sK);
            }
            for (int o = lane; o < NA * ND; o += 64) {
                const int a = o / ND, d = o % ND;
                double s = 0.0;
                for (int b = 0; b < NB; ++b)
                    for (int c = 0; c < NC; ++c) s += Dbc[b * NC + c] * VAT(a, b, c, d);
                fix_add_global(Kacc, (P.bfa + a) * nbf + Q.bfb + d, s, sK);
            }
            for (int o = lane; o < NB * ND; o += 64) {
                const int b = o / ND, d = o % ND;
                double s = 0.0;
                for (int a = 0; a < NA; ++a)
                    for (int c = 0; c < NC; ++c) s += Dac[a * NC + c] * VAT(a, b, c, d);
                fix_add_global(Kacc, (P.bfb + b) * nbf + Q.bfb + d, s, sK);
            }
        }
#undef VAT
    }
    if (do_j && lane < S::NAB && jp != 0.0) fix_add_global(Jacc, (P.bfa + lane / NB) * nbf + P.bfb + lane % NB, jp, sJ);
}

template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void jk_wave_kernel(
    const int4* __restrict__ items, const int2* __restrict__ segs, int nitems, const PairRec* __restrict__ bra, const PairRec* __restrict__ ket,
    int diag, const PrimPair<R>* __restrict__ pp, const R* __restrict__ rys_tab, const double* __restrict__ D,
    fix_t* __restrict__ Jacc, fix_t* __restrict__ Kacc, const double* __restrict__ fix, int do_j, int do_k,
    const int* __restrict__ ket_off, const int64_t* __restrict__ mat_off, const int* __restrict__ mol_nbf) {
    constexpr int WD = JkWave<LA, LB, LC, LD, R>::WAVE_D;
    __shared__ double smem[4 * WD];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int item = blockIdx.x * 4 + wid;
    if (item >= nitems) return;  // no block-wide barriers in the item: waves are independent
    jk_wave_item<LA, LB, LC, LD, R>(items[item], segs, lane, smem + wid * WD, bra, ket, diag, pp, rys_tab, D, Jacc, Kacc,
                                    fix, do_j, do_k, ket_off, mat_off, mol_nbf);
}

// Mixed precision in one launch: items [0, n64) FP64, the rest FP32 (one branch per wavefront).
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void jk_wave_kernel_mp(
    const int4* __restrict__ items, const int2* __restrict__ segs, int nitems, int n64, const PairRec* __restrict__ bra,
    const PairRec* __restrict__ ket, int diag, const PrimPair<double>* __restrict__ pp,
    const PrimPair<float>* __restrict__ ppf, const double* __restrict__ rys_tab, const float* __restrict__ rys_tabf,
    const double* __restrict__ D, fix_t* __restrict__ Jacc, fix_t* __restrict__ Kacc, const double* __restrict__ fix,
    int do_j, int do_k, const int* __restrict__ ket_off, const int64_t* __restrict__ mat_off,
    const int* __restrict__ mol_nbf) {
    constexpr int WD64 = JkWave<LA, LB, LC, LD, double>::WAVE_D, WD32 = JkWave<LA, LB, LC, LD, float>::WAVE_D;
    constexpr int WD = WD64 > WD32 ? WD64 : WD32;
    __shared__ double smem[4 * WD];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int item = blockIdx.x * 4 + wid;
    if (item >= nitems) return;
    if (item < n64)
        jk_wave_item<LA, LB, LC, LD, double>(items[item], segs, lane, smem + wid * WD, bra, ket, diag, pp, rys_tab, D, Jacc,
                                             Kacc, fix, do_j, do_k, ket_off, mat_off, mol_nbf);
    else
        jk_wave_item<LA, LB, LC, LD, float>(items[item], segs, lane, smem + wid * WD, bra, ket, diag, ppf, rys_tabf, D, Jacc,
                                            Kacc, fix, do_j, do_k, ket_off, mat_off, mol_nbf);
}

__device__ __forceinline__ int64_t pair_index64(int64_t i, int64_t j) {
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// The 8-fold compressed tensor (eri_compressed, integrals/eri/mod.rs:992-1070; index
// pair_index/eri_index :946-960).  Every element of a surviving quartet is written once.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void eri_tensor_kernel(
    const PairRec* __restrict__ bra, int nbra, const PairRec* __restrict__ ket, int nket, int diag,
    const PrimPair<double>* __restrict__ pp, const double* __restrict__ rys_tab, double* __restrict__ out,
    double thr) {
    using S = EriShape<LA, LB, LC, LD>;
    constexpr bool U = S::kUnroll;
    const int pidx = diag ? (nbra - 1 - (int)blockIdx.x) : (int)blockIdx.x;
    const PairRec P = bra[pidx];
    const int q1 = diag ? pidx + 1 : nket;  // screened below with the reference's predicate
    for (int q = threadIdx.x; q < q1; q += blockDim.x) {
        const PairRec Q = ket[q];
        if (P.Q * Q.Q < thr) continue;
        double eri[S::NQ];
        for_n<U, S::NQ>([&](int i) { eri[i] = 0.0; });
        contract_quartet<LA, LB, LC, LD, double>(P, Q, pp, rys_tab, eri);
#pragma unroll 1
        for (int idx = 0; idx < S::NQ; ++idx) {
            const int d = idx % S::ND, c = (idx / S::ND) % S::NC, b = (idx / S::NCD) % S::NB, a = idx / (S::NCD * S::NB);
            const int i = P.bfa + a, j = P.bfb + b;
            if (P.sa == P.sb && i < j) continue;
            const int k = Q.bfa + c, l = Q.bfb + d;
            if (Q.sa == Q.sb && k < l) continue;
            const int64_t ij = pair_index64(i, j), kl = pair_index64(k, l);
            if (diag && q == pidx && ij < kl) continue;
            const int64_t hi = ij > kl ? ij : kl, lo = ij > kl ? kl : ij;
            out[hi * (hi + 1) / 2 + lo] =
                eri[idx] * cnorm(LA, a) * cnorm(LB, b) * cnorm(LC, c) * cnorm(LD, d);
        }
    }
}

#define QCG_LA(c) ((c) == 0 ? 0 : (c) <= 2 ? 1 : 2)
#define QCG_LB(c) ((c) == 0 || (c) == 1 || (c) == 3 ? 0 : ((c) == 2 || (c) == 4) ? 1 : 2)

// Classes that have a per-lane kernel: up to kJkLaneMaxNQ Cartesian components (above that the
// wave-cooperative kernel always wins and the per-lane form is register-starved).
template <typename S>
inline constexpr bool kLaneForm = S::NQ <= kJkLaneMaxNQ;

template <int CB, int CK, typename R>
static hipError_t launch_jk_t(const JkLaunch& a, hipStream_t st) {
    constexpr int LA = QCG_LA(CB), LB = QCG_LB(CB), LC = QCG_LA(CK), LD = QCG_LB(CK);
    using S = EriShape<LA, LB, LC, LD>;
    if (a.nitems <= 0) return hipSuccess;
    // wave-cooperative mode (four items per block); classes without a per-lane form always take it
    if (a.block == 0 || !kLaneForm<S>) {
        if constexpr (JkWave<LA, LB, LC, LD, R>::kEnabled) {
            hipLaunchKernelGGL((jk_wave_kernel<LA, LB, LC, LD, R>), dim3((a.nitems + 3) / 4), dim3(256), 0, st,
                               a.items, a.segs, a.nitems, a.bra, a.ket, a.diag, (const PrimPair<R>*)a.pp, (const R*)a.rys,
                               a.D, a.Jacc, a.Kacc, a.fix, a.do_j, a.do_k, a.ket_off, a.mat_off, a.mol_nbf);
            return hipGetLastError();
        }
    }
    if constexpr (kLaneForm<S>) {
        // K and D rows of the bra functions in LDS when they fit (a.nbf is the largest molecule's nbf)
        size_t dyn_bytes = 0;
        const int bra_pp = QCG_JK_BRA_LDS ? a.bra_pp : 0;
        const int krow = jk_row_flags<LA, LB, LC, LD>(a.nbf, a.do_k, bra_pp, &dyn_bytes);
        auto go = [&](auto rows_c) {
            constexpr int ROWS = decltype(rows_c)::value;
            hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD, R, ROWS>), dim3(a.nitems), dim3(a.block > 0 ? a.block : 64),
                               dyn_bytes, st, a.items, a.segs, a.bra, a.ket, a.diag, (const PrimPair<R>*)a.pp,
                               (const R*)a.rys, a.D, a.Jacc, a.Kacc, a.fix, a.do_j, a.do_k, a.ket_off, a.mat_off,
                               a.mol_nbf, bra_pp);
        };
        if (krow == 3) go(std::integral_constant<int, 3>{});
        else if (krow == 1) go(std::integral_constant<int, 1>{});
        else go(std::integral_constant<int, 0>{});
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;  // not reached: large classes take the wave-cooperative form above
    }
}

// Mixed precision, both parts in one launch: a.items holds a.n64 FP64 items then the FP32 items,
// all cut for the same tile (a.block); a.pp/a.rys are the FP64 data, a.ppf/a.rysf the FP32 copies.
template <int CB, int CK>
static hipError_t launch_jk_mp_t(const JkLaunch& a, hipStream_t st) {
    constexpr int LA = QCG_LA(CB), LB = QCG_LB(CB), LC = QCG_LA(CK), LD = QCG_LB(CK);
    using S = EriShape<LA, LB, LC, LD>;
    if (a.nitems <= 0) return hipSuccess;
    // classes without a per-lane form always take the wave-cooperative one
    if (a.block == 0 || !kLaneForm<S>) {
        if constexpr (JkWave<LA, LB, LC, LD, double>::kEnabled) {
            hipLaunchKernelGGL((jk_wave_kernel_mp<LA, LB, LC, LD>), dim3((a.nitems + 3) / 4), dim3(256), 0, st,
                               a.items, a.segs, a.nitems, a.n64, a.bra, a.ket, a.diag, (const PrimPair<double>*)a.pp,
                               (const PrimPair<float>*)a.ppf, (const double*)a.rys, (const float*)a.rysf, a.D, a.Jacc,
                               a.Kacc, a.fix, a.do_j, a.do_k, a.ket_off, a.mat_off, a.mol_nbf);
            return hipGetLastError();
        }
    }
    if constexpr (kLaneForm<S>) {
        size_t dyn_bytes = 0;
        const int bra_pp = QCG_JK_BRA_LDS ? a.bra_pp : 0;
        const int krow = jk_row_flags<LA, LB, LC, LD>(a.nbf, a.do_k, bra_pp, &dyn_bytes);
        auto go = [&](auto rows_c) {
            constexpr int ROWS = decltype(rows_c)::value;
            hipLaunchKernelGGL((jk_kernel_mp<LA, LB, LC, LD, ROWS>), dim3(a.nitems), dim3(a.block > 0 ? a.block : 64),
                               dyn_bytes, st, a.items, a.segs, a.n64, a.bra, a.ket, a.diag,
                               (const PrimPair<double>*)a.pp, (const PrimPair<float>*)a.ppf, (const double*)a.rys,
                               (const float*)a.rysf, a.D, a.Jacc, a.Kacc, a.fix, a.do_j, a.do_k, a.ket_off, a.mat_off,
                               a.mol_nbf, bra_pp);
        };
        if (krow == 3) go(std::integral_constant<int, 3>{});
        else if (krow == 1) go(std::integral_constant<int, 1>{});
        else go(std::integral_constant<int, 0>{});
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;  // not reached: large classes take the wave-cooperative form above
    }
}

template <int CB, int CK>
static hipError_t launch_tensor_t(const JkLaunch& a, double* out, hipStream_t st) {
    constexpr int LA = QCG_LA(CB), LB = QCG_LB(CB), LC = QCG_LA(CK), LD = QCG_LB(CK);
    if (a.nbra <= 0 || a.nket <= 0) return hipSuccess;
    hipLaunchKernelGGL((eri_tensor_kernel<LA, LB, LC, LD>), dim3(a.nbra), dim3(256), 0, st, a.bra, a.nbra, a.ket,
                       a.nket, a.diag, (const PrimPair<double>*)a.pp, (const double*)a.rys, out, a.thr_lo);
    return hipGetLastError();
}

#define QCG_CAT2(a, b) a##b
#define QCG_CAT(a, b) QCG_CAT2(a, b)

hipError_t QCG_CAT(launch_jk_cb, QCG_CB)(int ck, bool fp32, const JkLaunch& a, hipStream_t st) {
    switch (ck) {
#define QCG_K(CK) \
    case CK: if constexpr (CK <= QCG_CB) return a.n64 >= 0 ? launch_jk_mp_t<QCG_CB, CK>(a, st) : fp32 ? launch_jk_t<QCG_CB, CK, float>(a, st) : launch_jk_t<QCG_CB, CK, double>(a, st); break;
        QCG_K(0) QCG_K(1) QCG_K(2) QCG_K(3) QCG_K(4) QCG_K(5)
#undef QCG_K
    }
    return hipErrorInvalidValue;
}

hipError_t QCG_CAT(launch_tensor_cb, QCG_CB)(int ck, const JkLaunch& a, double* out, hipStream_t st) {
    switch (ck) {
#define QCG_K(CK) \
    case CK: if constexpr (CK <= QCG_CB) return launch_tensor_t<QCG_CB, CK>(a, out, st); break;
        QCG_K(0) QCG_K(1) QCG_K(2) QCG_K(3) QCG_K(4) QCG_K(5)
#undef QCG_K
    }
    return hipErrorInvalidValue;
}

}  // namespace qcg
