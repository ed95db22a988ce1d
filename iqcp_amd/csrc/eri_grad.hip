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
// The density weight W is formed once per contracted quartet; the primitive/root loop then
// accumulates nine scalars (A, B, C x 3 directions): an adjoint contraction, no derivative
// integrals are ever stored.  Gradients of the bra atoms accumulate in registers, the ket atoms
// in LDS (natom x 3), flushed with FP64 atomics once per workgroup.
#include "eri_grad_device.h"
#include "eri_launch.h"

#ifndef QCG_CB
#error "compile with -DQCG_CB=<bra pair class>"
#endif

namespace qcg {

__device__ __forceinline__ int gfirst_below(const PairRec* __restrict__ ket, int qmax, double Qp, double thr) {
    int lo = 0, hi = qmax;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (Qp * ket[mid].Q >= thr) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void grad_kernel(const PairRec* __restrict__ bra, int nbra,
                                                   const PairRec* __restrict__ ket, int nket, int diag,
                                                   const PrimPair<double>* __restrict__ pp,
                                                   const double* __restrict__ rys_tab, const double* __restrict__ D,
                                                   int nbf, double xfac, double thr, int natom,
                                                   double* __restrict__ grad) {
    using S = EriShape<LA, LB, LC, LD>;
    // extended tables: powers up to (LA+1, LB+1, LC+1, LD)
    constexpr int EA = LA + 1, EB = LB + 1, EC = LC + 1, ED = LD;
    constexpr int NB_ = LA + LB + 2, NK_ = LC + LD + 1;
    constexpr int GE = (NB_ + 1) * (NK_ + 1);
    constexpr int HE = (EA + 1) * (EB + 1) * (EC + 1) * (ED + 1);
    constexpr int N = (NB_ + NK_) / 2 + 1;
    constexpr bool U = S::kUnroll;
    extern __shared__ double gl[];  // natom * 3 ket-atom gradient
    __shared__ int s_q1;
    const int pidx = diag ? (nbra - 1 - (int)blockIdx.x) : (int)blockIdx.x;
    const PairRec P = bra[pidx];
    const int tid = threadIdx.x;
    for (int i = tid; i < 3 * natom; i += blockDim.x) gl[i] = 0.0;
    if (tid == 0) s_q1 = gfirst_below(ket, diag ? pidx + 1 : nket, P.Q, thr);
    __syncthreads();
    const int q1 = s_q1;
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    const double two_pi_52 = 2.0 * kPi * kPi * 1.7724538509055160273;
    auto hidx = [](int i, int j, int k, int l) { return ((i * (EB + 1) + j) * (EC + 1) + k) * (ED + 1) + l; };
    for (int q = tid; q < q1; q += blockDim.x) {
        const PairRec Q = ket[q];
        if (P.Q * Q.Q < thr) continue;
        const double deg = ((P.sa == P.sb) ? 0.5 : 1.0) * ((Q.sa == Q.sb) ? 0.5 : 1.0) * ((diag && q == pidx) ? 0.5 : 1.0);
        double W[S::NQ];
        grad_weights<LA, LB, LC, LD>(P, Q, D, nbf, xfac, deg, W);
        double gq[9];
        grad_quartet<LA, LB, LC, LD>(P, Q, pp, rys_tab, W, gq);
        // centre D = -(A + B + C)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            gA[x] += gq[x];
            gB[x] += gq[3 + x];
            atomicAdd(&gl[3 * Q.atom_a + x], gq[6 + x]);
            atomicAdd(&gl[3 * Q.atom_b + x], -(gq[x] + gq[3 + x] + gq[6 + x]));
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        double a = gA[x], b = gB[x];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            b += __shfl_xor(b, off, 64);
        }
        if ((tid & 63) == 0) {
            atomicAdd(&gl[3 * P.atom_a + x], a);
            atomicAdd(&gl[3 * P.atom_b + x], b);
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
    if (a.nbra <= 0 || a.nket <= 0) return hipSuccess;
    const size_t lds = 3 * (size_t)a.natom * sizeof(double);
    hipLaunchKernelGGL((grad_kernel<LA, LB, LC, LD>), dim3(a.nbra), dim3(256), lds, st, a.bra, a.nbra, a.ket, a.nket,
                       a.diag, a.pp, a.rys, a.D, a.nbf, a.xfac, a.thr, a.natom, a.grad);
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
