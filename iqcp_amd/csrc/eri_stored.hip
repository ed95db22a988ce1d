// Fock build from a stored compressed two-electron tensor (build_fock, scf/mod.rs:825-881): the
// preset-system path, where S, H and the 8-fold compressed tensor come from a file instead of a
// basis.  One wavefront per lower-triangle element (mu >= nu); for each lambda the lanes take
// sigma <= lambda, with the reference's grouping of the terms:
//   lambda == sigma: D_ll [(mn|ll) - 1/2 (ml|nl)]
//   sigma  <  lambda: D_ls [2 (mn|ls) - 1/2 ((ml|ns) + (ms|nl))]
// then a fixed-order wavefront reduction, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "stored_launch.h"

namespace qcg {

namespace {

__device__ __forceinline__ int64_t pair_ix(int64_t i, int64_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

__device__ __forceinline__ int64_t eri_ix(int i, int j, int k, int l) {
    const int64_t p = pair_ix(i, j), q = pair_ix(k, l);
    return pair_ix(p, q);
}

}  // namespace

__global__ __launch_bounds__(256) void fock_stored_kernel(int nbf, const double* __restrict__ H,
                                                          const double* __restrict__ D,
                                                          const double* __restrict__ eri, double* __restrict__ F) {
    const int64_t w = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (w >= (int64_t)nbf * (nbf + 1) / 2) return;  // the whole wavefront leaves together
    int mu = (int)((sqrt(8.0 * (double)w + 1.0) - 1.0) * 0.5);
    while ((int64_t)mu * (mu + 1) / 2 > w) --mu;
    while ((int64_t)(mu + 1) * (mu + 2) / 2 <= w) ++mu;
    const int nu = (int)(w - (int64_t)mu * (mu + 1) / 2);
    double g = 0.0;
    for (int la = 0; la < nbf; ++la)
        for (int s0 = 0; s0 <= la; s0 += 64) {
            const int si = s0 + lane;
            if (si < la) {
                const double j = eri[eri_ix(mu, nu, la, si)];
                const double k1 = eri[eri_ix(mu, la, nu, si)];
                const double k2 = eri[eri_ix(mu, si, nu, la)];
                g += D[(int64_t)la * nbf + si] * (2.0 * j - 0.5 * (k1 + k2));
            } else if (si == la) {
                g += D[(int64_t)la * nbf + la] * (eri[eri_ix(mu, nu, la, la)] - 0.5 * eri[eri_ix(mu, la, nu, la)]);
            }
        }
    for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off, 64);
    if (lane == 0) {
        F[(int64_t)mu * nbf + nu] = H[(int64_t)mu * nbf + nu] + g;
        if (mu != nu) F[(int64_t)nu * nbf + mu] = H[(int64_t)nu * nbf + mu] + g;
    }
}

hipError_t launch_fock_stored(int nbf, const double* H, const double* D, const double* eri, double* F,
                              hipStream_t st) {
    const int64_t nel = (int64_t)nbf * (nbf + 1) / 2;
    if (nbf <= 0) return hipSuccess;
    if ((nel + 3) / 4 > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fock_stored_kernel, dim3((unsigned)((nel + 3) / 4)), dim3(256), 0, st, nbf, H, D, eri, F);
    return hipGetLastError();
}

}  // namespace qcg
