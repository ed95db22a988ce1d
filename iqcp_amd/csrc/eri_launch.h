// Host-side launch interface of the two-electron kernels.
#pragma once
#include <algorithm>
#include <vector>

#include "fixacc.h"
#include "qcg_types.h"

namespace qcg {

// Work items over ket segments.  The surviving kets of one bra pair are one contiguous range per
// contraction bucket; the host lists those ranges as segments (first ket, count) and cuts their
// concatenation into items (bra index, first segment, offset into it, w), so a block's lanes stay
// busy across bucket boundaries instead of idling at the end of every bucket.
//
// w packs the ket count (bits 0-15) and an optional range of the bra's primitive pairs: first
// (bits 16-23) and count (bits 24-31, 0 = all).  A bra pair with many primitive pairs is cut into
// such ranges so that no lane runs the whole npp_bra x npp_ket primitive loop of a heavy quartet
// alone; the integral is linear in the bra primitives, so the partial quartets are digested
// (J and K) like whole ones.
constexpr int kItemMaxKets = 0xffff;
constexpr int kItemMaxPrim = 0xff;

// items of at most ch kets over segments [g0, segs.size()) of bra `bra`, primitive range
// (pbeg, pcnt) (pcnt = 0: all of them)
inline void cut_items(std::vector<int4>& items, const std::vector<int2>& segs, int g0, int bra, int ch,
                      int pbeg = 0, int pcnt = 0) {
    int total = 0;
    for (size_t g = (size_t)g0; g < segs.size(); ++g) total += segs[g].y;
    const int wbits = (pbeg << 16) | (pcnt << 24);
    int g = g0, r = 0;
    for (int off = 0; off < total; off += ch) {
        const int n = std::min(ch, total - off);
        items.push_back(make_int4(bra, g, r, n | wbits));
        r += n;
        while (g < (int)segs.size() && r >= segs[g].y) r -= segs[g++].y;
    }
}

// cut_items for every primitive range of a bra pair with npp primitive pairs: ranges of at most
// split pairs (split <= 0 or npp <= split: one range, the whole pair), sizes balanced
inline void cut_items_split(std::vector<int4>& items, const std::vector<int2>& segs, int g0, int bra, int ch,
                            int npp, int split) {
    if (split <= 0 || npp <= split || npp > kItemMaxPrim) {
        cut_items(items, segs, g0, bra, ch);
        return;
    }
    const int nr = (npp + split - 1) / split;
    for (int k = 0; k < nr; ++k) {
        const int b = k * npp / nr, e = (k + 1) * npp / nr;
        cut_items(items, segs, g0, bra, ch, b, e - b);
    }
}

inline QCG_HD int item_kets(const int4& it) { return it.w & kItemMaxKets; }

// Most bra primitive pairs one per-lane J/K item may have: the kernel stages them in LDS (a block's
// bra pair is uniform, so the inner primitive loop reads them as LDS broadcasts instead of scalar
// loads).  The planner cuts heavier bras into primitive ranges of at most this many.
constexpr int kBraPPMax = 64;

// Largest class (Cartesian components) with a per-lane J/K kernel; larger classes always run the
// wave-cooperative one, classes from 27 components up to this limit may run either (tile block 0).
#ifndef QCG_JK_LANE_MAX_NQ
#define QCG_JK_LANE_MAX_NQ 81
#endif
constexpr int kJkLaneMaxNQ = QCG_JK_LANE_MAX_NQ;

// the bra pair of an item, restricted to the item's primitive range
inline QCG_HD PairRec item_bra(const PairRec* __restrict__ bra, const int4& it) {
    PairRec P = bra[it.x];
    const int pcnt = (it.w >> 24) & kItemMaxPrim;
    if (pcnt) {
        P.pp_off += (it.w >> 16) & kItemMaxPrim;
        P.npp = pcnt;
    }
    return P;
}

// Device cursor over an item's kets: position r counts kets from the start of segment g; at()
// returns the ket index and may be called only for the item's own kets (j < item count), which
// keeps the walk inside the bra's segments.
struct KetCursor {
    const int2* __restrict__ seg;
    int g, r;
    __device__ __forceinline__ int at() {
        int2 s = seg[g];
        while (r >= s.y) {
            r -= s.y;
            s = seg[++g];
        }
        return s.x + r;
    }
};

struct JkLaunch {
    const PairRec* bra;
    int nbra;
    const PairRec* ket;
    int nket;
    int diag;           // bra class == ket class: only ket positions <= bra position
    const void* pp;     // PrimPair<double> or PrimPair<float>
    const void* rys;    // Rys table (double or float)
    const double* D;    // density, nbf x nbf, symmetric
    int nbf;
    fix_t* Jacc;        // fixed-point accumulators (fixacc.h), one word per element
    fix_t* Kacc;
    const double* fix;  // per molecule: scale of J, scale of K, and their inverses (fix_scale_kernel)
    double thr_lo;      // process quartets with thr_lo <= Q_P Q_Q < thr_hi
    double thr_hi;
    int do_j, do_k;
    // batch of molecules (nmol = 1 for a single molecule): kets of bra pair P are
    // ket[ket_off[P.mol] .. + ket_cnt[P.mol]); matrices of molecule m start at mat_off[m],
    // dimension mol_nbf[m]; nbf above is the largest mol_nbf.
    const int* ket_off;
    const int* ket_cnt;
    const int64_t* mat_off;
    const int* mol_nbf;
    // J/K work items (index in the bra class list, first segment, offset into it, ket count) over
    // the ket segments segs[g] = (first ket, count), ket indices relative to the molecule's list
    // (see KetCursor); one block of `block` threads per item
    const int4* items;
    const int2* segs;
    int nitems;
    int block;
    int bra_pp;         // most primitive pairs of any bra of the items (<= kBraPPMax): LDS staging
    // mixed precision in one launch (n64 >= 0): the first n64 items are FP64, the rest FP32 with
    // the FP32 primitive pairs ppf and Rys table rysf; n64 < 0: one precision, given by the caller
    int n64 = -1;
    const void* ppf = nullptr;
    const void* rysf = nullptr;
};

hipError_t launch_jk_cb0(int ck, bool fp32, const JkLaunch& a, hipStream_t st);
hipError_t launch_jk_cb1(int ck, bool fp32, const JkLaunch& a, hipStream_t st);
hipError_t launch_jk_cb2(int ck, bool fp32, const JkLaunch& a, hipStream_t st);
hipError_t launch_jk_cb3(int ck, bool fp32, const JkLaunch& a, hipStream_t st);
hipError_t launch_jk_cb4(int ck, bool fp32, const JkLaunch& a, hipStream_t st);
hipError_t launch_jk_cb5(int ck, bool fp32, const JkLaunch& a, hipStream_t st);
hipError_t launch_tensor_cb0(int ck, const JkLaunch& a, double* out, hipStream_t st);
hipError_t launch_tensor_cb1(int ck, const JkLaunch& a, double* out, hipStream_t st);
hipError_t launch_tensor_cb2(int ck, const JkLaunch& a, double* out, hipStream_t st);
hipError_t launch_tensor_cb3(int ck, const JkLaunch& a, double* out, hipStream_t st);
hipError_t launch_tensor_cb4(int ck, const JkLaunch& a, double* out, hipStream_t st);
hipError_t launch_tensor_cb5(int ck, const JkLaunch& a, double* out, hipStream_t st);

inline hipError_t launch_jk(int cb, int ck, bool fp32, const JkLaunch& a, hipStream_t st) {
    switch (cb) {
        case 0: return launch_jk_cb0(ck, fp32, a, st);
        case 1: return launch_jk_cb1(ck, fp32, a, st);
        case 2: return launch_jk_cb2(ck, fp32, a, st);
        case 3: return launch_jk_cb3(ck, fp32, a, st);
        case 4: return launch_jk_cb4(ck, fp32, a, st);
        case 5: return launch_jk_cb5(ck, fp32, a, st);
    }
    return hipErrorInvalidValue;
}

inline hipError_t launch_eri_tensor(int cb, int ck, const JkLaunch& a, double* out, hipStream_t st) {
    switch (cb) {
        case 0: return launch_tensor_cb0(ck, a, out, st);
        case 1: return launch_tensor_cb1(ck, a, out, st);
        case 2: return launch_tensor_cb2(ck, a, out, st);
        case 3: return launch_tensor_cb3(ck, a, out, st);
        case 4: return launch_tensor_cb4(ck, a, out, st);
        case 5: return launch_tensor_cb5(ck, a, out, st);
    }
    return hipErrorInvalidValue;
}

struct GradLaunch {
    const PairRec* bra;
    int nbra;
    const PairRec* ket;
    int nket;
    int diag;
    const PrimPair<double>* pp;  // ea and p - ea are the two primitive exponents
    const double* rys;
    const double* D;
    int nbf;
    double xfac;  // exchange weight c = hf_frac / 4 (eri_gradient_fused, scf/gradient.rs:1438)
    double thr;
    int natom;
    double* grad;  // 3 * natom, accumulated
    // work items (index in the bra class list, first segment, offset into it, ket count) over the
    // ket segments segs (see KetCursor): one wavefront, or one block of lanes, each
    const int4* items;
    const int2* segs;
    int nitems;
};

// Work-item length (kets per item) the gradient kernel of a class expects: one block of threads
// per item for the light classes (per-lane kernel), a few kets per wavefront otherwise.
inline int grad_item_kets(int cb, int ck) {
    auto nc = [](int c) { return c == 0 ? 1 : c == 1 ? 3 : c == 2 ? 9 : c == 3 ? 6 : c == 4 ? 18 : 36; };
    return nc(cb) * nc(ck) <= 54 ? 128 : 8;
}

hipError_t launch_grad_cb0(int ck, const GradLaunch& a, hipStream_t st);
hipError_t launch_grad_cb1(int ck, const GradLaunch& a, hipStream_t st);
hipError_t launch_grad_cb2(int ck, const GradLaunch& a, hipStream_t st);
hipError_t launch_grad_cb3(int ck, const GradLaunch& a, hipStream_t st);
hipError_t launch_grad_cb4(int ck, const GradLaunch& a, hipStream_t st);
hipError_t launch_grad_cb5(int ck, const GradLaunch& a, hipStream_t st);

inline hipError_t launch_grad(int cb, int ck, const GradLaunch& a, hipStream_t st) {
    switch (cb) {
        case 0: return launch_grad_cb0(ck, a, st);
        case 1: return launch_grad_cb1(ck, a, st);
        case 2: return launch_grad_cb2(ck, a, st);
        case 3: return launch_grad_cb3(ck, a, st);
        case 4: return launch_grad_cb4(ck, a, st);
        case 5: return launch_grad_cb5(ck, a, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_schwarz(int c, const PairRec* pairs, int n, const PrimPair<double>* pp, const double* rys,
                          double* Q, hipStream_t st);
hipError_t launch_fill(double* x, size_t n, double v, hipStream_t st);
// What the fixed-point scale kernels read (device pointers, packed per molecule): the Schwarz bounds
// of the shell pairs as nshell x nshell matrices (qsh, offsets qsh_off), their maximum (qmax), the
// shell of every function (bf_shell, offsets bf_off into the functions of all molecules), the first
// function of every shell plus a sentinel (shell_first, offsets shell_first_off), the molecule of
// every function (row_mol, nrows functions in all) and 2 doubles of scratch per function (rowsum).
struct FixData {
    const int64_t* mat_off;
    const int* mol_nbf;
    const double* qsh;
    const int64_t* qsh_off;
    const int* mol_ns;
    const int* bf_shell;
    const int* bf_off;
    const int* row_mol;
    int nrows;
    double* rowsum;
    const double* qmax;
    const int* shell_first;
    const int* shell_first_off;
};
// Per molecule m the fixed-point scales fix[4m] (J) and fix[4m + 1] (K) and their inverses from a
// bound on the accumulated magnitudes at density D: B_J = 2 Qmax sum_ls Q_ls |D_ls|,
// B_K = 2 Qmax max_s sum_t Q_st sum_(n in t) sum_l |D_nl|.  Also zeroes the accumulators Jacc and
// Kacc (either may be null).
hipError_t launch_fix_scale(const double* D, const FixData& f, double* fix, int nmol, fix_t* Jacc, fix_t* Kacc,
                            hipStream_t st);
hipError_t launch_symmetrize(const fix_t* Jacc, const fix_t* Kacc, const double* fix, double* J, double* K, int nbf,
                             hipStream_t st);
// batched J = Jacc + Jacc^T, K = Kacc + Kacc^T over block-diagonal packed matrices
hipError_t launch_symmetrize_batch(const fix_t* Jacc, const fix_t* Kacc, const double* fix, double* J, double* K, int nmol,
                                   const int64_t* mat_off, const int* mol_nbf, int max_nbf, hipStream_t st);
hipError_t launch_fock_combine(const double* H, const fix_t* Jacc, const fix_t* Kacc, const double* fix, double* F, int nbf,
                               double kfac, hipStream_t st);
hipError_t launch_boys_batch(int m, int n, const double* T, double* F, int* status, hipStream_t st);
hipError_t launch_rys_batch(int nroots, int n, const double* T, const double* tab, double* x, double* w,
                            hipStream_t st);
hipError_t launch_rys_batch_f32(int nroots, int n, const float* T, const float* tab, float* x, float* w,
                                hipStream_t st);

}  // namespace qcg
