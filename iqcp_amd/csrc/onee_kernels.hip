// One-electron integrals and gradient terms on the device.
//
// Gradient: one thread per (ordered shell pair, primitive pair) and term (blockIdx.y = 0: overlap
// and kinetic; blockIdx.y = 1 + c: attraction to nucleus c), per-atom sums in LDS.
#include "onee_pair.h"
#include "onee_launch.h"

namespace qcg {

__global__ __launch_bounds__(128) void onee_grad_kernel(BasisView B, const int4* __restrict__ work, int nwork,
                                                       const double* __restrict__ D, const double* __restrict__ W,
                                                       const double* __restrict__ rys_tab, double* __restrict__ grad) {
    extern __shared__ double gl[];  // natom x 3
    const int n3 = 3 * B.natom;
    for (int i = threadIdx.x; i < n3; i += blockDim.x) gl[i] = 0.0;
    __syncthreads();
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nwork) {
        const int4 w = work[u];
        auto add = [&](int atom, int x, double v) {
            if (v != 0.0) __hip_atomic_fetch_add(&gl[3 * atom + x], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        if (blockIdx.y == 0) onee_grad_overlap_kinetic(B, w.x, w.y, w.z, w.w, D, W, add);
        else onee_grad_nuclear(B, w.x, w.y, w.z, w.w, (int)blockIdx.y - 1, D, rys_tab, add);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n3; i += blockDim.x)
        if (gl[i] != 0.0) atomicAdd(&grad[i], gl[i]);
}

// S, T, V: one wavefront per Cartesian component pair of the shell pairs si <= sj (work = {si, sj,
// ia, jb}); its lanes take the (primitive pair, term) combinations, term = overlap+kinetic or one
// nucleus, and a fixed-order wavefront reduction sums them.  The wavefront writes its element
// and, off the diagonal blocks, the mirrored one.
__global__ __launch_bounds__(256) void onee_matrix_kernel(BasisView B, const int4* __restrict__ work, int nwork,
                                                         const double* __restrict__ rys_tab, double* __restrict__ S,
                                                         double* __restrict__ T, double* __restrict__ V) {
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (u >= nwork) return;  // the whole wavefront leaves together
    const int4 w = work[u];
    const int nb = B.nprim[w.y], nterm = B.natom + 1;
    const int ncomb = B.nprim[w.x] * nb * nterm;
    double out[3] = {0.0, 0.0, 0.0};
    for (int t = lane; t < ncomb; t += 64) {
        const int pp = t / nterm, c = t - pp * nterm;
        onee_element_term(B, w.x, w.y, w.z, w.w, pp / nb, pp % nb, c == B.natom ? -1 : c, rys_tab, out);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) out[k] += __shfl_xor(out[k], off, 64);
    if (lane != 0) return;
    const double nrm = onee_norm_comp(B.l[w.x], w.z) * onee_norm_comp(B.l[w.y], w.w);
    const int nbf = B.nbf;
    const size_t i = (size_t)B.bf_off[w.x] + w.z, j = (size_t)B.bf_off[w.y] + w.w;
    S[i * nbf + j] = nrm * out[0];
    T[i * nbf + j] = nrm * out[1];
    V[i * nbf + j] = nrm * out[2];
    if (w.x != w.y) {
        S[j * nbf + i] = nrm * out[0];
        T[j * nbf + i] = nrm * out[1];
        V[j * nbf + i] = nrm * out[2];
    }
}

hipError_t launch_onee_matrices(const BasisView& B, const int4* work, int nwork, const double* rys_tab, double* S,
                                double* T, double* V, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(onee_matrix_kernel, dim3((nwork + 3) / 4), dim3(256), 0, st, B, work, nwork, rys_tab, S, T, V);
    return hipGetLastError();
}

hipError_t launch_onee_grad(const BasisView& B, const int4* work, int nwork, const double* D, const double* W,
                            const double* rys_tab, double* grad, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    const size_t lds = 3 * (size_t)B.natom * sizeof(double);
    if (lds > 65536) return hipErrorInvalidValue;
    if (B.natom + 1 > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(onee_grad_kernel, dim3((nwork + 127) / 128, B.natom + 1), dim3(128), lds, st, B, work, nwork, D,
                       W, rys_tab, grad);
    return hipGetLastError();
}

}  // namespace qcg
