// Grid contractions of the XC build on the matrix cores (MFMA), element type T = float (mixed
// precision) or double:
//   xc_rho_kernel : X = Phi D one 64-point block at a time on MFMA, reduced on the fly into rho and
//                   grad rho (X never goes to HBM), then B3LYP per point (FP64), the aow rows and
//                   E_xc -- the work of compute_density_and_gradient_matmul and the point part of
//                   build_vxc_gga_matmul (dft/ks_scf.rs:1028-1093, :1150-1296) in one kernel;
//   xc_v_kernel   : V = Phi^T aow as 64x64 output tiles over grid chunks (split K), one partial
//                   matrix per chunk, summed and symmetrised afterwards (sum_parts, sym_half).
// rho, grad rho, E_xc and the functional are accumulated in FP64 whatever T is; with T = float the
// GEMM products are FP32 MFMA with FP32 accumulation.
#include "xc_functional.h"
#include "xc_launch.h"

namespace qcg {

typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x4_t __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64: D (16x16) += A (16x4) B (4x16).  Lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[row(l / 16, r)][l % 16] in element r, with
// row = 4 (l / 16) + r for FP32 and (l / 16) + 4 r for FP64 (both checked on the device).
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using acc = f32x4_t;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};
template <>
struct Mfma<double> {
    using acc = f64x4_t;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};

constexpr int kRhoBM = 64;  // grid points per block, 16 per wavefront
constexpr int kRhoBK = 16;  // basis functions per K step
constexpr int kPad = 16;    // LDS row padding: the 16 lanes of a fragment row hit distinct banks
constexpr int kVBM = 64;    // V output tile (64 x 64), 16 rows per wavefront
constexpr int kVBK = 16;    // grid points per K step

// NT: MFMA column tiles (16 columns each) of X per pass over the columns; a pass covers 16 NT
// columns, all passes together the nbf columns of X.
template <typename T, int NT>
__global__ __launch_bounds__(256) void xc_rho_kernel(int G, int nbf, const double* __restrict__ w,
                                                     const T* __restrict__ phi, const T* __restrict__ dpx,
                                                     const T* __restrict__ dpy, const T* __restrict__ dpz,
                                                     const T* __restrict__ Dm, T* __restrict__ aow,
                                                     double* __restrict__ exc_sum) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int BN = 16 * NT;
    constexpr int NA = kRhoBM * kRhoBK / 256, NB = kRhoBK * BN / 256;  // staged elements per thread
    static_assert(kRhoBM * kRhoBK % 256 == 0 && kRhoBK * BN % 256 == 0, "tile sizes");
    __shared__ T As[kRhoBK][kRhoBM + kPad];  // Phi block, transposed: As[k][point]
    __shared__ T Bs[kRhoBK][BN + kPad];      // rows k0.. of D, this pass's columns
    __shared__ double red[4][16][4];         // per wavefront and point: rho, d rho / dx, dy, dz sums
    __shared__ double coef[4][16][4];        // per wavefront and point: w vrho, 4 w vsigma grad rho
    __shared__ double part[4];
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
    const int g0 = blockIdx.x * kRhoBM;
    double rs[4] = {0.0, 0.0, 0.0, 0.0}, xs[4] = {0.0, 0.0, 0.0, 0.0}, ys[4] = {0.0, 0.0, 0.0, 0.0},
           zs[4] = {0.0, 0.0, 0.0, 0.0};
    T ra[NA], rb[NB];  // next K step's tiles, loaded while the current one is multiplied
    auto load = [&](int n0, int k0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = tid + 256 * i, r = e / kRhoBK, c = e % kRhoBK, g = g0 + r, k = k0 + c;
            ra[i] = (g < G && k < nbf) ? phi[(size_t)g * nbf + k] : T(0);
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = tid + 256 * i, r = e / BN, c = e % BN, k = k0 + r, n = n0 + c;
            rb[i] = (k < nbf && n < nbf) ? Dm[(size_t)k * nbf + n] : T(0);
        }
    };
    for (int n0 = 0; n0 < nbf; n0 += BN) {
        acc_t acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = acc_t{T(0), T(0), T(0), T(0)};
        load(n0, 0);
        for (int k0 = 0; k0 < nbf; k0 += kRhoBK) {
            __syncthreads();  // the previous step's fragment reads are done
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                const int e = tid + 256 * i;
                As[e % kRhoBK][e / kRhoBK] = ra[i];
            }
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int e = tid + 256 * i;
                Bs[e / BN][e % BN] = rb[i];
            }
            __syncthreads();
            if (k0 + kRhoBK < nbf) load(n0, k0 + kRhoBK);
#pragma unroll
            for (int ks = 0; ks < kRhoBK; ks += 4) {
                const T a = As[ks + lk][16 * wv + li];
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t] = M::mma(a, Bs[ks + lk][16 * t + li], acc[t]);
            }
        }
        // rho = sum_mu X Phi, grad rho / 2 = sum_mu X grad Phi over this pass's columns
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int col = n0 + 16 * t + li;
            if (col < nbf) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int g = g0 + 16 * wv + M::row(lk, r);
                    if (g < G) {
                        const size_t idx = (size_t)g * nbf + col;
                        const double x = (double)acc[t][r];
                        rs[r] += x * (double)phi[idx];
                        xs[r] += x * (double)dpx[idx];
                        ys[r] += x * (double)dpy[idx];
                        zs[r] += x * (double)dpz[idx];
                    }
                }
            }
            // a bounded number of tiles' loads in flight: hoisting all of them costs hundreds of VGPRs
            if (t % 2 == 1) asm volatile("" ::: "memory");
        }
    }
    // sums over the 16 lanes that share a point row (same l / 16)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            rs[r] += __shfl_xor(rs[r], off, 64);
            xs[r] += __shfl_xor(xs[r], off, 64);
            ys[r] += __shfl_xor(ys[r], off, 64);
            zs[r] += __shfl_xor(zs[r], off, 64);
        }
    }
    if (li == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            red[wv][M::row(lk, r)][0] = rs[r];
            red[wv][M::row(lk, r)][1] = xs[r];
            red[wv][M::row(lk, r)][2] = ys[r];
            red[wv][M::row(lk, r)][3] = zs[r];
        }
    }
    __syncthreads();
    double contrib = 0.0;
    if (l < 16) {
        const int g = g0 + 16 * wv + l;
        double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
        if (g < G) {
            const double rho = fmax(red[wv][l][0], 0.0);
            const double grx = 2.0 * red[wv][l][1], gry = 2.0 * red[wv][l][2], grz = 2.0 * red[wv][l][3];
            double exc, vrho, vsig;
            xc_b3lyp(rho, grx * grx + gry * gry + grz * grz, exc, vrho, vsig);
            const double wg = w[g], f4 = 4.0 * wg * vsig;
            c0 = wg * vrho;
            c1 = f4 * grx;
            c2 = f4 * gry;
            c3 = f4 * grz;
            contrib = wg * exc * rho;
        }
        coef[wv][l][0] = c0;
        coef[wv][l][1] = c1;
        coef[wv][l][2] = c2;
        coef[wv][l][3] = c3;
    }
    __syncthreads();
    // aow = w vrho Phi + 4 w vsigma (grad rho . grad Phi), row by row of the wavefront's points
    for (int p = 0; p < 16; ++p) {
        const int g = g0 + 16 * wv + p;
        if (g >= G) break;  // uniform over the wavefront
        const double c0 = coef[wv][p][0], c1 = coef[wv][p][1], c2 = coef[wv][p][2], c3 = coef[wv][p][3];
        const size_t row = (size_t)g * nbf;
        for (int col = l; col < nbf; col += 64) {
            const size_t idx = row + col;
            aow[idx] = (T)(c0 * (double)phi[idx] + c1 * (double)dpx[idx] + c2 * (double)dpy[idx] + c3 * (double)dpz[idx]);
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) contrib += __shfl_xor(contrib, off, 64);
    if (l == 0) part[wv] = contrib;
    __syncthreads();
    if (tid == 0) {
        const double s = part[0] + part[1] + part[2] + part[3];
        if (s != 0.0) atomicAdd(exc_sum, s);
    }
}

// Vparts[chunk] = Phi[chunk rows]^T aow[chunk rows]; block -> (chunk, tile) so that the tiles of one
// chunk share an XCD (blocks are dealt round-robin over the 8 XCDs) and reuse its Phi and aow rows
// from that XCD's L2.
template <typename T>
__global__ __launch_bounds__(256) void xc_v_kernel(int G, int nbf, int ntile, int nchunk, int chunk,
                                                   const T* __restrict__ phi, const T* __restrict__ aow,
                                                   double* __restrict__ Vparts) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    __shared__ T As[kVBK][kVBM + kPad];  // Phi[g][m0 + i]
    __shared__ T Bs[kVBK][kVBM + kPad];  // aow[g][n0 + j]
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nt2 = ntile * ntile;
    const int ch = (slot / nt2) * 8 + xcd, tile = slot % nt2;
    if (ch >= nchunk) return;  // the whole block: before any barrier
    const int m0 = (tile / ntile) * kVBM, n0 = (tile % ntile) * kVBM;
    const int gb = ch * chunk, ge = min(G, gb + chunk);
    acc_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = acc_t{T(0), T(0), T(0), T(0)};
    constexpr int NE = kVBK * kVBM / 256;  // staged elements per thread and operand
    T ra[NE], rb[NE];                       // next K step's rows, loaded during the current MFMAs
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i, r = e / kVBM, c = e % kVBM, g = k0 + r;
            const bool in = g < ge;
            ra[i] = (in && m0 + c < nbf) ? phi[(size_t)g * nbf + m0 + c] : T(0);
            rb[i] = (in && n0 + c < nbf) ? aow[(size_t)g * nbf + n0 + c] : T(0);
        }
    };
    load(gb);
    for (int k0 = gb; k0 < ge; k0 += kVBK) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            As[e / kVBM][e % kVBM] = ra[i];
            Bs[e / kVBM][e % kVBM] = rb[i];
        }
        __syncthreads();
        if (k0 + kVBK < ge) load(k0 + kVBK);
#pragma unroll
        for (int ks = 0; ks < kVBK; ks += 4) {
            const T a = As[ks + lk][16 * wv + li];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = M::mma(a, Bs[ks + lk][16 * t + li], acc[t]);
        }
    }
    double* out = Vparts + (size_t)ch * nbf * nbf;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int col = n0 + 16 * t + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = m0 + 16 * wv + M::row(lk, r);
            if (row < nbf && col < nbf) out[(size_t)row * nbf + col] = (double)acc[t][r];
        }
    }
}

// FP64 correction of E_xc for the rounding of D to FP32: dE_xc/dD = V_xc, so
// E_xc(D) = E_xc(D32) + sum V_xc (D - D32) + O(|D - D32|^2) with |D - D32| ~ 6e-8 |D|.
__global__ __launch_bounds__(256) void exc_d32_correction_kernel(const double* __restrict__ V,
                                                                 const double* __restrict__ D,
                                                                 const float* __restrict__ D32, size_t n,
                                                                 double* __restrict__ exc_sum) {
    __shared__ double part[4];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += V[i] * (D[i] - (double)D32[i]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = part[0] + part[1] + part[2] + part[3];
        if (t != 0.0) atomicAdd(exc_sum, t);
    }
}

hipError_t launch_exc_d32_correction(const double* V, const double* D, const float* D32, size_t n, double* exc_sum,
                                     hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>(256, (n + 255) / 256);
    hipLaunchKernelGGL(exc_d32_correction_kernel, dim3(blocks), dim3(256), 0, st, V, D, D32, n, exc_sum);
    return hipGetLastError();
}

__global__ void to_f32_kernel(const double* __restrict__ in, float* __restrict__ out, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

hipError_t launch_to_f32(const double* in, float* out, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_xc_mfma_t(int G, int nbf, int nchunk, const double* w, const T* phi, const T* dpx,
                                   const T* dpy, const T* dpz, const T* Dm, T* aow, double* exc_sum, double* Vparts,
                                   hipStream_t st) {
    if (G <= 0 || nbf <= 0 || nchunk <= 0) return hipErrorInvalidValue;
    // column passes of at most 8 MFMA tiles, with the tiles spread evenly over the passes
    const int tiles = (nbf + 15) / 16, passes = (tiles + 7) / 8, nt = (tiles + passes - 1) / passes;
    const dim3 grid((G + kRhoBM - 1) / kRhoBM);
    switch (nt) {
#define QCG_RHO_NT(n)                                                                                   \
    case n:                                                                                             \
        hipLaunchKernelGGL((xc_rho_kernel<T, n>), grid, dim3(256), 0, st, G, nbf, w, phi, dpx, dpy, dpz, Dm, \
                           aow, exc_sum);                                                               \
        break;
        QCG_RHO_NT(1) QCG_RHO_NT(2) QCG_RHO_NT(3) QCG_RHO_NT(4) QCG_RHO_NT(5) QCG_RHO_NT(6) QCG_RHO_NT(7) QCG_RHO_NT(8)
#undef QCG_RHO_NT
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int chunk = (G + nchunk - 1) / nchunk;
    const int nch = (G + chunk - 1) / chunk;  // <= nchunk partial matrices written
    const int ntile = (nbf + kVBM - 1) / kVBM;
    const int nblk = 8 * ((nch + 7) / 8) * ntile * ntile;
    hipLaunchKernelGGL((xc_v_kernel<T>), dim3(nblk), dim3(256), 0, st, G, nbf, ntile, nch, chunk, phi, aow, Vparts);
    return hipGetLastError();
}

hipError_t launch_xc_mfma(int G, int nbf, int nchunk, const double* w, const float* phi, const float* dpx,
                          const float* dpy, const float* dpz, const float* Dm, float* aow, double* exc_sum,
                          double* Vparts, hipStream_t st) {
    return launch_xc_mfma_t<float>(G, nbf, nchunk, w, phi, dpx, dpy, dpz, Dm, aow, exc_sum, Vparts, st);
}

hipError_t launch_xc_mfma(int G, int nbf, int nchunk, const double* w, const double* phi, const double* dpx,
                          const double* dpy, const double* dpz, const double* Dm, double* aow, double* exc_sum,
                          double* Vparts, hipStream_t st) {
    return launch_xc_mfma_t<double>(G, nbf, nchunk, w, phi, dpx, dpy, dpz, Dm, aow, exc_sum, Vparts, st);
}

}  // namespace qcg
