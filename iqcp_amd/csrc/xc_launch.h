// Host-side launch interface of the XC kernels.
#pragma once
#include "qcg_common.h"

namespace qcg {

struct GridShell {
    double cx, cy, cz, amin;  // centre, smallest exponent (screening)
    int l, nprim, poff, bf;
    int atom, pad;
};

// Becke partition weights of np points (owner[p]: the point's atom), on the device
hipError_t launch_becke_partition(int64_t np, const double* pts, const int* owner, int nat, const double* axyz,
                                  const double* dist, double* part, hipStream_t st);
hipError_t launch_basis_on_grid(int G, int nshell, const GridShell* shells, const double* exps, const double* cn,
                                const double* pts, int nbf, double* phi, double* dphx, double* dphy, double* dphz,
                                hipStream_t st);
hipError_t launch_xc_points(int G, int nbf, const double* w, const double* phi, const double* dphx,
                            const double* dphy, const double* dphz, const double* X, double* aow, double* exc_sum,
                            double* rho_out, hipStream_t st);
hipError_t launch_xc_eval(int n, const double* rho, const double* sigma, double* exc, double* vrho, double* vsigma,
                          hipStream_t st);
hipError_t launch_sym_half(const double* V, double* out, int n, hipStream_t st);
// MFMA form of the XC build (xc_mfma.hip): E_xc added to exc_sum, aow rows written, and the
// partial products V_k = Phi[chunk k]^T aow[chunk k] in Vparts for k < xc_mfma_chunks(G, nchunk)
// (at most nchunk); the caller sums them and symmetrises.
// FP32: phi for the GEMM operands, g4 = (Phi, dPhi/dx, dPhi/dy, dPhi/dz) interleaved per element
hipError_t launch_xc_mfma(int G, int nbf, int nchunk, const double* w, const float* phi, const float4* g4,
                          const float* Dm, float* aow, double* exc_sum, double* Vparts, hipStream_t st);
hipError_t launch_xc_mfma(int G, int nbf, int nchunk, const double* w, const double* phi, const double* dpx,
                          const double* dpy, const double* dpz, const double* Dm, double* aow, double* exc_sum,
                          double* Vparts, hipStream_t st);
// Output tile width of the MFMA V kernel in 16-wide MFMA tiles per wavefront (VW; a block is
// 2 x 2 wavefronts, a 32 VW square): the width whose tiles pad the nbf x nbf matrix least, the
// larger one on ties.  nbf = 304 (19 MFMA tiles): VW = 5, 2 x 2 tiles of 160 covering 320^2
// instead of 3 x 3 tiles of 128 covering 384^2 (in FP64 the 25 accumulators of a wavefront take
// 200 AGPRs, one wavefront per SIMD, and it still wins).
inline int xc_v_width(int nbf, bool /*fp32*/) {
    const int n16 = (nbf + 15) / 16;
    int best = 4, bpad = -1;
    for (int w = 5; w >= 3; --w) {
        const int t = (n16 + 2 * w - 1) / (2 * w) * 2 * w;
        if (bpad < 0 || t * t < bpad) {
            bpad = t * t;
            best = w;
        }
    }
    return best;
}
// output tiles of the MFMA V kernel over the nbf x nbf matrix
inline int xc_mfma_vtiles(int nbf, bool fp32) {
    const int b = 32 * xc_v_width(nbf, fp32);
    const int t = (nbf + b - 1) / b;
    return t * t;
}
// Blocks of the MFMA V kernel (FP32 or FP64 products) resident per CU at this basis size's tile
// width, from the runtime's occupancy query
int xc_v_blocks_per_cu(int nbf, bool fp32);
inline int xc_mfma_chunks(int G, int nchunk) {
    const int chunk = (G + nchunk - 1) / nchunk;
    return (G + chunk - 1) / chunk;
}
hipError_t launch_to_f32(const double* in, float* out, size_t n, hipStream_t st);
// out[i] = (a[i], b[i], c[i], d[i]) in FP32
hipError_t launch_to_f32x4(const double* a, const double* b, const double* c, const double* d, float4* out, size_t n,
                           hipStream_t st);
// exc_sum += sum_i V[i] (D[i] - D32[i]): the first-order FP64 correction of E_xc for the FP32
// rounding of the density matrix (V = V_xc of the same build)
hipError_t launch_exc_d32_correction(const double* V, const double* D, const float* D32, size_t n, double* exc_sum,
                                     hipStream_t st);
// out[i] = sum_k parts[k n + i]
hipError_t launch_sum_parts(const double* parts, int nparts, size_t n, double* out, hipStream_t st);
// XC nuclear gradient: per-point A rows and t = 2 w vsigma grad rho, then the per-(shell, point)
// accumulation into grad (natom x 3) from X = Phi D and M = A D.
hipError_t launch_xc_grad_points(int G, int nbf, const double* w, const double* phi, const double* dphx,
                                 const double* dphy, const double* dphz, const double* X, double* A, double* t,
                                 hipStream_t st);
hipError_t launch_xc_grad_accumulate(int G, int nshell, const GridShell* shells, const double* exps, const double* cn,
                                     const double* pts, int nbf, const double* X, const double* M, const double* t,
                                     double* grad, hipStream_t st);

}  // namespace qcg
