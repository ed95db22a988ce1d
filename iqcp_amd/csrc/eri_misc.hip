// Small kernels around the J/K build: Schwarz bounds, Boys/Rys batches, fills, J/K symmetrisation.
#include "boys.h"
#include "eri_device.h"
#include "eri_launch.h"

namespace qcg {

// Schwarz bound per pair: sqrt(max |(ab|ab)|) (compute_schwarz_bounds, integrals/eri/mod.rs:1263-1290)
template <int LA, int LB>
__global__ __launch_bounds__(128) void schwarz_kernel(const PairRec* __restrict__ pairs, int n,
                                                      const PrimPair<double>* __restrict__ pp,
                                                      const double* __restrict__ rys_tab, double* __restrict__ Q) {
    using S = EriShape<LA, LB, LA, LB>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PairRec P = pairs[i];
    double eri[S::NQ];
    for_n<S::kUnroll, S::NQ>([&](int k) { eri[k] = 0.0; });
    contract_quartet<LA, LB, LA, LB, double>(P, P, pp, rys_tab, eri);
    double mx = 0.0;
#pragma unroll 1
    for (int k = 0; k < S::NQ; ++k) {
        const int d = k % S::NB, c = (k / S::NB) % S::NA, b = (k / S::NAB) % S::NB, a = k / (S::NAB * S::NB);
        const double nrm = cnorm(LA, a) * cnorm(LB, b) * cnorm(LA, c) * cnorm(LB, d);
        mx = fmax(mx, fabs(eri[k] * nrm));
    }
    Q[i] = sqrt(mx);
}

hipError_t launch_schwarz(int c, const PairRec* pairs, int n, const PrimPair<double>* pp, const double* rys,
                          double* Q, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    dim3 g((n + 127) / 128), b(128);
    switch (c) {
        case 0: hipLaunchKernelGGL((schwarz_kernel<0, 0>), g, b, 0, st, pairs, n, pp, rys, Q); break;
        case 1: hipLaunchKernelGGL((schwarz_kernel<1, 0>), g, b, 0, st, pairs, n, pp, rys, Q); break;
        case 2: hipLaunchKernelGGL((schwarz_kernel<1, 1>), g, b, 0, st, pairs, n, pp, rys, Q); break;
        case 3: hipLaunchKernelGGL((schwarz_kernel<2, 0>), g, b, 0, st, pairs, n, pp, rys, Q); break;
        case 4: hipLaunchKernelGGL((schwarz_kernel<2, 1>), g, b, 0, st, pairs, n, pp, rys, Q); break;
        case 5: hipLaunchKernelGGL((schwarz_kernel<2, 2>), g, b, 0, st, pairs, n, pp, rys, Q); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void fill_kernel(double* __restrict__ x, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = v;
}

hipError_t launch_fill(double* x, size_t n, double v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, n, v);
    return hipGetLastError();
}

// Fixed-point scales of one build (fixacc.h).  Every J contribution to element (m, n) is a product
// (mn|ls) D_ls with |(mn|ls)| <= Q_mn Q_ls (Schwarz, per component), so
// |Jacc_mn| <= Qmax sum_ls Q_ls |D_ls| (counted twice for the symmetry images: B_J); a K contribution
// to (m, l) is (mn|ls) D_ns, so |Kacc_ml| <= Qmax sum_t Q_sh(m),t rho_t with rho_t = sum over the
// functions n of shell t of sum_s |D_ns| (twice: B_K).  Two kernels: one wavefront per row of D for
// the row sums, then one block per molecule for the shell sums and the scales.  Every reduction runs
// in a fixed order, so the scales, like the sums they serve, are the same on every build of the same
// density.
__global__ __launch_bounds__(256) void fix_rows_kernel(
    const double* __restrict__ D, const int64_t* __restrict__ mat_off, const int* __restrict__ mol_nbf,
    const double* __restrict__ qsh, const int64_t* __restrict__ qsh_off, const int* __restrict__ mol_ns,
    const int* __restrict__ bf_shell, const int* __restrict__ bf_off, const int* __restrict__ row_mol, int nrows,
    double* __restrict__ rowsum, fix_t* __restrict__ Jacc, fix_t* __restrict__ Kacc) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= nrows) return;
    const int m = row_mol[row], n = row - bf_off[m], nbf = mol_nbf[m], ns = mol_ns[m];
    const int* bs = bf_shell + bf_off[m];
    const size_t ro = mat_off[m] + (size_t)n * nbf;
    const double* Drow = D + ro;
    const double* Qrow = qsh + qsh_off[m] + (size_t)bs[n] * ns;
    double r = 0.0, j = 0.0;
    for (int l = lane; l < nbf; l += 64) {
        const double a = fabs(Drow[l]);
        r += a;
        j += Qrow[bs[l]] * a;
        // the same row of the accumulators starts at zero (null: that matrix is not built)
        if (Jacc) Jacc[ro + l] = 0;
        if (Kacc) Kacc[ro + l] = 0;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        r += __shfl_xor(r, off, 64);
        j += __shfl_xor(j, off, 64);
    }
    if (lane == 0) {
        rowsum[2 * row] = r;
        rowsum[2 * row + 1] = j;
    }
}

constexpr int kFixMaxShells = 4096;  // shells per molecule the scale kernel handles (rho in LDS)

__global__ __launch_bounds__(1024) void fix_scale_kernel(FixData f, double* __restrict__ fix) {
    __shared__ double red[1024];
    __shared__ double rho[kFixMaxShells];
    const int m = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbf = f.mol_nbf[m], ns = f.mol_ns[m];
    const double* Q = f.qsh + f.qsh_off[m];
    const int* first = f.shell_first + f.shell_first_off[m];
    const double* rs = f.rowsum + 2 * (size_t)f.bf_off[m];
    // shell row sums rho_t over the shell's consecutive functions, and sum_n jrow_n
    double aj = 0.0;
    for (int n = tid; n < nbf; n += blockDim.x) aj += rs[2 * n + 1];
    for (int t = tid; t < ns; t += blockDim.x) {
        double v = 0.0;
        for (int n = first[t]; n < first[t + 1]; ++n) v += rs[2 * n];
        rho[t] = v;
    }
    red[tid] = aj;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (tid < h) red[tid] += red[tid + h];
        __syncthreads();
    }
    const double sumJ = red[0];
    __syncthreads();
    // max over shells s of sum_t Q_st rho_t: one wavefront per shell s, lanes over t
    double km = 0.0;
    for (int sh = wave; sh < ns; sh += 16) {
        double t = 0.0;
        for (int u = lane; u < ns; u += 64) t += Q[(size_t)sh * ns + u] * rho[u];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off, 64);
        km = fmax(km, t);
    }
    red[tid] = km;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (tid < h) red[tid] = fmax(red[tid], red[tid + h]);
        __syncthreads();
    }
    if (tid == 0) {
        const double qmax = f.qmax[m];
        const double sj = fix_scale_for(2.0 * qmax * sumJ), sk = fix_scale_for(2.0 * qmax * red[0]);
        fix[4 * m] = sj;
        fix[4 * m + 1] = sk;
        fix[4 * m + 2] = 1.0 / sj;
        fix[4 * m + 3] = 1.0 / sk;
    }
}

hipError_t launch_fix_scale(const double* D, const FixData& f, double* fix, int nmol, fix_t* Jacc, fix_t* Kacc,
                            hipStream_t st) {
    hipLaunchKernelGGL(fix_rows_kernel, dim3((f.nrows + 3) / 4), dim3(256), 0, st, D, f.mat_off, f.mol_nbf, f.qsh,
                       f.qsh_off, f.mol_ns, f.bf_shell, f.bf_off, f.row_mol, f.nrows, f.rowsum, Jacc, Kacc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fix_scale_kernel, dim3(nmol), dim3(1024), 0, st, f, fix);
    return hipGetLastError();
}

// Element (i, j) of A + A^T for a fixed-point accumulator: the integer words are summed before
// the conversion, so the result is exactly symmetric.
__device__ __forceinline__ double fix_sym(const fix_t* __restrict__ a, size_t ij, size_t ji, double inv) {
    return fix_value(a[ij] + a[ji], inv);
}

// J = Jacc + Jacc^T, K = Kacc + Kacc^T
__global__ void symmetrize_kernel(const fix_t* __restrict__ Jacc, const fix_t* __restrict__ Kacc,
                                  const double* __restrict__ fix, double* __restrict__ J, double* __restrict__ K,
                                  int nbf) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nbf) return;
    const size_t ij = (size_t)i * nbf + j, ji = (size_t)j * nbf + i;
    if (J) J[ij] = fix_sym(Jacc, ij, ji, fix[2]);
    if (K) K[ij] = fix_sym(Kacc, ij, ji, fix[3]);
}

__global__ void symmetrize_batch_kernel(const fix_t* __restrict__ Jacc, const fix_t* __restrict__ Kacc,
                                        const double* __restrict__ fix, double* __restrict__ J,
                                        double* __restrict__ K, const int64_t* __restrict__ mat_off,
                                        const int* __restrict__ mol_nbf) {
    const int m = blockIdx.z, n = mol_nbf[m];
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || j >= n) return;
    const size_t o = (size_t)mat_off[m];
    const size_t ij = o + (size_t)i * n + j, ji = o + (size_t)j * n + i;
    if (J) J[ij] = fix_sym(Jacc, ij, ji, fix[4 * m + 2]);
    if (K) K[ij] = fix_sym(Kacc, ij, ji, fix[4 * m + 3]);
}

hipError_t launch_symmetrize_batch(const fix_t* Jacc, const fix_t* Kacc, const double* fix, double* J, double* K,
                                   int nmol, const int64_t* mat_off, const int* mol_nbf, int max_nbf, hipStream_t st) {
    hipLaunchKernelGGL(symmetrize_batch_kernel, dim3((max_nbf + 127) / 128, max_nbf, nmol), dim3(128), 0, st, Jacc,
                       Kacc, fix, J, K, mat_off, mol_nbf);
    return hipGetLastError();
}

hipError_t launch_symmetrize(const fix_t* Jacc, const fix_t* Kacc, const double* fix, double* J, double* K, int nbf,
                             hipStream_t st) {
    hipLaunchKernelGGL(symmetrize_kernel, dim3((nbf + 127) / 128, nbf), dim3(128), 0, st, Jacc, Kacc, fix, J, K, nbf);
    return hipGetLastError();
}

// F = H + J - kfac K  (build_fock, scf/mod.rs:825-881 with kfac = 1/2; hybrid form dft/ks_scf.rs:1960)
__global__ void fock_combine_kernel(const double* __restrict__ H, const fix_t* __restrict__ Jacc,
                                    const fix_t* __restrict__ Kacc, const double* __restrict__ fix,
                                    double* __restrict__ F, int nbf, double kfac) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nbf) return;
    const size_t ij = (size_t)i * nbf + j, ji = (size_t)j * nbf + i;
    F[ij] = H[ij] + fix_sym(Jacc, ij, ji, fix[2]) - kfac * fix_sym(Kacc, ij, ji, fix[3]);
}

hipError_t launch_fock_combine(const double* H, const fix_t* Jacc, const fix_t* Kacc, const double* fix, double* F,
                               int nbf, double kfac, hipStream_t st) {
    hipLaunchKernelGGL(fock_combine_kernel, dim3((nbf + 127) / 128, nbf), dim3(128), 0, st, H, Jacc, Kacc, fix, F, nbf,
                       kfac);
    return hipGetLastError();
}

// Batched Boys F_0..F_m (boys_eval_all, boys/mod.rs:327): one T per lane.
__global__ void boys_batch_kernel(int m, int n, const double* __restrict__ T, double* __restrict__ F,
                                  int* __restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f[51];
    const int rc = boys_all(m, T[i], f);
    if (rc) {
        atomicOr(status, 1);
        for (int k = 0; k <= m; ++k) F[(size_t)i * (m + 1) + k] = 0.0;
        return;
    }
    for (int k = 0; k <= m; ++k) F[(size_t)i * (m + 1) + k] = f[k];
}

hipError_t launch_boys_batch(int m, int n, const double* T, double* F, int* status, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(boys_batch_kernel, dim3((n + 255) / 256), dim3(256), 0, st, m, n, T, F, status);
    return hipGetLastError();
}

// Batched Rys nodes/weights (rys_roots, rys/mod.rs:253): one T per lane.
template <typename R>
__global__ void rys_batch_kernel(int nroots, int n, const R* __restrict__ T, const R* __restrict__ tab,
                                 R* __restrict__ x, R* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    R xr[QCG_RYS_NMAX], wr[QCG_RYS_NMAX];
    rys_eval_n<R>(nroots, T[i], tab, xr, wr);
    for (int k = 0; k < nroots; ++k) {
        x[(size_t)i * nroots + k] = xr[k];
        w[(size_t)i * nroots + k] = wr[k];
    }
}

hipError_t launch_rys_batch(int nroots, int n, const double* T, const double* tab, double* x, double* w,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((rys_batch_kernel<double>), dim3((n + 255) / 256), dim3(256), 0, st, nroots, n, T, tab, x, w);
    return hipGetLastError();
}

hipError_t launch_rys_batch_f32(int nroots, int n, const float* T, const float* tab, float* x, float* w,
                                hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((rys_batch_kernel<float>), dim3((n + 255) / 256), dim3(256), 0, st, nroots, n, T, tab, x, w);
    return hipGetLastError();
}

}  // namespace qcg
