// Grid contractions of the XC build on the matrix cores (MFMA), element type T = float (mixed
// precision) or double:
//   xc_rho_kernel : X = Phi D one 64-point block at a time on MFMA, reduced on the fly into rho and
//                   grad rho (X never goes to HBM), then B3LYP per point (FP64), the aow rows and
//                   E_xc -- the work of compute_density_and_gradient_matmul and the point part of
//                   build_vxc_gga_matmul (dft/ks_scf.rs:1028-1093, :1150-1296) in one kernel;
//   xc_rho_all_kernel : the same for nbf <= 352 with X for all columns held per wavefront, so Phi
//                   is read once per K step instead of once per column pass;
//   xc_v_kernel   : V = Phi^T aow as 128x128 output tiles over grid chunks (split K), one partial
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

constexpr int kRhoMT = 1;                 // MFMA row tiles (16 points each) per wavefront
constexpr int kRhoBM = 4 * 16 * kRhoMT;   // grid points per block
constexpr int kRhoPW = 16 * kRhoMT;       // grid points per wavefront
constexpr int kRhoBK = 16;                // basis functions per K step
constexpr int kPad = 16;                  // LDS row padding: the lanes of a fragment row hit distinct banks
// V: VW MFMA tiles per wavefront and dimension (xc_v_width, 3 to 5), so the block's output tile
// (2 x 2 wavefronts) is 32 VW square
constexpr int kVBK = 16;                  // grid points per K step

// NT: MFMA column tiles (16 columns each) of X per pass over the columns; a pass covers 16 NT
// columns, all passes together the nbf columns of X.  Each wavefront owns kRhoPW points (kRhoMT
// row tiles), so every D fragment read from LDS feeds kRhoMT MFMAs.
// Phi, dPhi/dx, dPhi/dy, dPhi/dz at (point g, function mu) = element idx = g nbf + mu: FP32 from
// one interleaved float4 array (one 16-byte load), FP64 from the four separate arrays.
template <typename T>
struct GridVals {
    const T *phi, *dpx, *dpy, *dpz;
    const float4* g4;
    __device__ __forceinline__ void get(size_t idx, double& f, double& x, double& y, double& z) const {
        if constexpr (sizeof(T) == 4) {
            const float4 v = g4[idx];
            f = v.x; x = v.y; y = v.z; z = v.w;
        } else {
            f = phi[idx]; x = dpx[idx]; y = dpy[idx]; z = dpz[idx];
        }
    }
};

// LDS row padding that makes a fragment read conflict-free: the four rows l / 16 of one read start
// 16 elements apart modulo the bank width (32 doubles or 64 floats), so with 16 consecutive
// elements per row the 64 lanes cover the banks evenly.
template <typename T>
constexpr int lds_pad(int n) {
    return sizeof(T) == 8 ? (16 - n % 32 + 32) % 32 : (16 - n % 64 + 64) % 64;
}

// All-columns rho kernel options.  PREFETCH_D: the next K step's D rows are loaded into registers
// during the current step's MFMAs (FP32: on; FP64: off, the registers cost the second block per CU:
// (H2O)16 FP64 rho 11.2 -> 9.3 ms without).  WPE64: waves per SIMD the kernel is compiled for (4
// = two 512-thread blocks per CU, at most 128 VGPRs).
#ifndef QCG_XC_RHO_PREFETCH_D
#define QCG_XC_RHO_PREFETCH_D 1
#endif
#ifndef QCG_XC_RHO_PREFETCH_D64
#define QCG_XC_RHO_PREFETCH_D64 0
#endif
#ifndef QCG_XC_RHO_WPE64
#define QCG_XC_RHO_WPE64 4
#endif

// Most MFMA column tiles the all-columns kernel holds (nbf <= 352): D rows of one K step, all
// columns, stay within the 64 KB of static LDS per block.
constexpr int kRhoAllMaxNT = 22;

// The point part shared by both rho kernels, once the rho and grad rho sums of the block's 64 points
// are in red[row tile][point]: B3LYP per point (FP64), the aow rows, E_xc.  NW wavefronts per block
// (4 or 8); wavefront wv works on row tile wv % 4, and with 8 wavefronts each takes half its rows
// in the aow pass.
template <typename T, int NW, int MAXC = 0>
__device__ __forceinline__ void xc_point_tail(int G, int nbf, int wv, int l, int g0, const double* __restrict__ w,
                                              const GridVals<T>& gv, T* __restrict__ aow, double* __restrict__ exc_sum,
                                              double (*red)[kRhoPW][4], double (*coef)[kRhoPW][4], double* part) {
    static_assert(NW == 4 || NW == 8, "four or eight wavefronts per block");
    const int tid = 64 * wv + l, rt = wv & 3, gw = g0 + kRhoPW * rt;
    double contrib = 0.0;
    if (wv < 4 && l < kRhoPW) {
        const int g = gw + l;
        double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
        if (g < G) {
            const double rho = fmax(red[rt][l][0], 0.0);
            const double grx = 2.0 * red[rt][l][1], gry = 2.0 * red[rt][l][2], grz = 2.0 * red[rt][l][3];
            double exc, vrho, vsig;
            xc_b3lyp(rho, grx * grx + gry * gry + grz * grz, exc, vrho, vsig);
            const double wg = w[g], f4 = 4.0 * wg * vsig;
            c0 = wg * vrho;
            c1 = f4 * grx;
            c2 = f4 * gry;
            c3 = f4 * grz;
            contrib = wg * exc * rho;
        }
        coef[rt][l][0] = c0;
        coef[rt][l][1] = c1;
        coef[rt][l][2] = c2;
        coef[rt][l][3] = c3;
    }
    __syncthreads();
    // aow = w vrho Phi + 4 w vsigma (grad rho . grad Phi), row by row of the wavefront's points
    constexpr int kRows = kRhoPW / (NW / 4);
    const int p0 = (wv >> 2) * kRows;
    for (int p = p0; p < p0 + kRows; ++p) {
        const int g = gw + p;
        if (g >= G) break;  // uniform over the wavefront
        const double c0 = coef[rt][p][0], c1 = coef[rt][p][1], c2 = coef[rt][p][2], c3 = coef[rt][p][3];
        const size_t row = (size_t)g * nbf;
        if constexpr (MAXC > 0) {
            // nbf <= 64 MAXC: the row's loads all in flight before the first use (unconditional, at
            // columns clamped into the basis, so none waits inside a branch of its own)
            double f[MAXC], dx[MAXC], dy[MAXC], dz[MAXC];
#pragma unroll
            for (int i = 0; i < MAXC; ++i) gv.get(row + min(l + 64 * i, nbf - 1), f[i], dx[i], dy[i], dz[i]);
#pragma unroll
            for (int i = 0; i < MAXC; ++i)
                if (l + 64 * i < nbf) aow[row + l + 64 * i] = (T)(c0 * f[i] + c1 * dx[i] + c2 * dy[i] + c3 * dz[i]);
        } else {
            for (int col = l; col < nbf; col += 64) {
                double f, dx, dy, dz;
                gv.get(row + col, f, dx, dy, dz);
                aow[row + col] = (T)(c0 * f + c1 * dx + c2 * dy + c3 * dz);
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) contrib += __shfl_xor(contrib, off, 64);
    if (l == 0 && wv < 4) part[wv] = contrib;
    __syncthreads();
    if (tid == 0) {
        const double s = part[0] + part[1] + part[2] + part[3];
        if (s != 0.0) atomicAdd(exc_sum, s);
    }
}

template <typename T, int NT>
__global__ __launch_bounds__(256) void xc_rho_kernel(int G, int nbf, const double* __restrict__ w,
                                                     const T* __restrict__ phi, const GridVals<T> gv,
                                                     const T* __restrict__ Dm, T* __restrict__ aow,
                                                     double* __restrict__ exc_sum) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int BN = 16 * NT;
    constexpr int NA = kRhoBM * kRhoBK / 256, NB = kRhoBK * BN / 256;  // staged elements per thread
    static_assert(kRhoBM * kRhoBK % 256 == 0 && kRhoBK * BN % 256 == 0, "tile sizes");
    __shared__ T As[kRhoBK][kRhoBM + kPad];  // Phi block, transposed: As[k][point]
    __shared__ T Bs[kRhoBK][BN + kPad];      // rows k0.. of D, this pass's columns
    __shared__ double red[4][kRhoPW][4];     // per wavefront and point: rho, d rho / dx, dy, dz sums
    __shared__ double coef[4][kRhoPW][4];    // per wavefront and point: w vrho, 4 w vsigma grad rho
    __shared__ double part[4];
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
    const int g0 = blockIdx.x * kRhoBM, gw = g0 + kRhoPW * wv;
    double rs[kRhoMT][4], xs[kRhoMT][4], ys[kRhoMT][4], zs[kRhoMT][4];
#pragma unroll
    for (int m = 0; m < kRhoMT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) rs[m][r] = xs[m][r] = ys[m][r] = zs[m][r] = 0.0;
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
        acc_t acc[kRhoMT][NT];
#pragma unroll
        for (int m = 0; m < kRhoMT; ++m)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[m][t] = acc_t{T(0), T(0), T(0), T(0)};
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
                T a[kRhoMT];
#pragma unroll
                for (int m = 0; m < kRhoMT; ++m) a[m] = As[ks + lk][kRhoPW * wv + 16 * m + li];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const T b = Bs[ks + lk][16 * t + li];
#pragma unroll
                    for (int m = 0; m < kRhoMT; ++m) acc[m][t] = M::mma(a[m], b, acc[m][t]);
                }
            }
        }
        // rho = sum_mu X Phi, grad rho / 2 = sum_mu X grad Phi over this pass's columns
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int col = n0 + 16 * t + li;
            if (col < nbf) {
#pragma unroll
                for (int m = 0; m < kRhoMT; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int g = gw + 16 * m + M::row(lk, r);
                        if (g < G) {
                            double f, dx, dy, dz;
                            gv.get((size_t)g * nbf + col, f, dx, dy, dz);
                            const double x = (double)acc[m][t][r];
                            rs[m][r] += x * f;
                            xs[m][r] += x * dx;
                            ys[m][r] += x * dy;
                            zs[m][r] += x * dz;
                        }
                    }
            }
            // one tile's loads in flight at a time: hoisting all of them costs hundreds of VGPRs
            asm volatile("" ::: "memory");
        }
    }
    // sums over the 16 lanes that share a point row (same l / 16)
#pragma unroll
    for (int m = 0; m < kRhoMT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                rs[m][r] += __shfl_xor(rs[m][r], off, 64);
                xs[m][r] += __shfl_xor(xs[m][r], off, 64);
                ys[m][r] += __shfl_xor(ys[m][r], off, 64);
                zs[m][r] += __shfl_xor(zs[m][r], off, 64);
            }
        }
    if (li == 0) {
#pragma unroll
        for (int m = 0; m < kRhoMT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = 16 * m + M::row(lk, r);
                red[wv][p][0] = rs[m][r];
                red[wv][p][1] = xs[m][r];
                red[wv][p][2] = ys[m][r];
                red[wv][p][3] = zs[m][r];
            }
    }
    __syncthreads();
    xc_point_tail<T, 4>(G, nbf, wv, l, g0, w, gv, aow, exc_sum, red, coef, part);
}

#if QCG_XC_RHO_WPE64
#define QCG_XC_RHO_ATTR __attribute__((amdgpu_waves_per_eu(QCG_XC_RHO_WPE64)))
#else
#define QCG_XC_RHO_ATTR
#endif

// All-columns form of xc_rho_kernel for nbf <= 16 kRhoAllMaxNT, eight wavefronts per block of 64
// points: wavefront wv keeps the X tile of the 16 points of row tile wv % 4 for column half wv / 4
// (NH = ceil(NT / 2) MFMA column tiles, one accumulator each), so Phi is read from HBM once per K
// step instead of once per column pass; D rows k0..k0+15 (all columns) are staged in LDS per K step
// (D stays L2-resident: every block reads all of it).
template <typename T, int NT>
__global__ __launch_bounds__(512) QCG_XC_RHO_ATTR void xc_rho_all_kernel(int G, int nbf, const double* __restrict__ w,
                                                         const T* __restrict__ phi, const GridVals<T> gv,
                                                         const T* __restrict__ Dm, T* __restrict__ aow,
                                                         double* __restrict__ exc_sum) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int NH = (NT + 1) / 2, BN = 16 * NT, SA = kRhoBM + lds_pad<T>(kRhoBM);
    // Bs row stride: padded for conflict-free fragment reads, and Bs large enough to hold red2 after
    // the K loop (2 x 4 x 16 x 4 doubles)
    constexpr int SB0 = BN + lds_pad<T>(BN), SBM = 4096 / (kRhoBK * (int)sizeof(T));
    constexpr int SB = SB0 >= SBM ? SB0 : SBM;
    constexpr int NA = kRhoBM * kRhoBK / 512;  // Phi elements staged per thread and K step
    static_assert(kRhoMT == 1 && kRhoBM == 64 && NA * 512 == kRhoBM * kRhoBK, "one 16-point row tile per wavefront");
    __shared__ T As[kRhoBK][SA];  // Phi block, transposed: As[k][point]
    __shared__ T Bs[kRhoBK][SB];  // D rows k0 .. k0 + 15, all columns
    __shared__ double red[4][kRhoPW][4];
    __shared__ double coef[4][kRhoPW][4];
    __shared__ double part[4];
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
    const int rt = wv & 3, half = wv >> 2, t0 = half * NH;
    const int g0 = blockIdx.x * kRhoBM, gw = g0 + kRhoPW * rt;
    acc_t acc[NH];
#pragma unroll
    for (int t = 0; t < NH; ++t) acc[t] = acc_t{T(0), T(0), T(0), T(0)};
    // next K step's Phi (and, with QCG_XC_RHO_PREFETCH_D, D) elements, loaded while the current step
    // is multiplied
    constexpr bool kPrefD = sizeof(T) == 8 ? QCG_XC_RHO_PREFETCH_D64 : QCG_XC_RHO_PREFETCH_D;
    constexpr int NBD = kPrefD ? (kRhoBK * BN + 511) / 512 : 1;
    T ra[NA], rb[NBD];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = tid + 512 * i, r = e / kRhoBK, c = e % kRhoBK, g = g0 + r, k = k0 + c;
            ra[i] = (g < G && k < nbf) ? phi[(size_t)g * nbf + k] : T(0);
        }
        if constexpr (kPrefD) {
#pragma unroll
            for (int i = 0; i < NBD; ++i) {
                const int e = tid + 512 * i, r = e / BN, c = e - r * BN, k = k0 + r;
                rb[i] = (e < kRhoBK * BN && k < nbf && c < nbf) ? Dm[(size_t)k * nbf + c] : T(0);
            }
        }
    };
    load(0);
    for (int k0 = 0; k0 < nbf; k0 += kRhoBK) {
        __syncthreads();  // the previous step's fragment reads are done
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = tid + 512 * i;
            As[e % kRhoBK][e / kRhoBK] = ra[i];
        }
        if constexpr (kPrefD) {
#pragma unroll
            for (int i = 0; i < NBD; ++i) {
                const int e = tid + 512 * i, r = e / BN, c = e - r * BN;
                if (e < kRhoBK * BN) Bs[r][c] = rb[i];
            }
        } else {
            for (int e = tid; e < kRhoBK * BN; e += 512) {
                const int r = e / BN, c = e - r * BN, k = k0 + r;
                Bs[r][c] = (k < nbf && c < nbf) ? Dm[(size_t)k * nbf + c] : T(0);
            }
        }
        __syncthreads();
        if (k0 + kRhoBK < nbf) load(k0 + kRhoBK);
        // two k sub-steps per unrolled body: all four hoist 4 NH fragment reads (spills at NH = 8)
#pragma unroll 2
        for (int ks = 0; ks < kRhoBK; ks += 4) {
            const T a = As[ks + lk][kRhoPW * rt + li];
#pragma unroll
            for (int t = 0; t < NH; ++t)
                if (t0 + t < NT) acc[t] = M::mma(a, Bs[ks + lk][16 * (t0 + t) + li], acc[t]);
        }
    }
    // rho = sum_mu X Phi, grad rho / 2 = sum_mu X grad Phi over this half's columns
    double rs[4] = {}, xs[4] = {}, ys[4] = {}, zs[4] = {};
#pragma unroll
    for (int t = 0; t < NH; ++t) {
        // the tile's loads are unconditional, at addresses clamped into the grid and the basis (X is
        // zero outside them, and the product is masked too), so the four rows' loads issue back to
        // back instead of each waiting inside its own branch: (H2O)16 XC, mixed 5.78 -> 5.41 ms,
        // FP64 products 14.25 -> 12.49 ms
        const int col = 16 * (t0 + t) + li;
        const bool cin = col < nbf;
        const int cc = cin ? col : nbf - 1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int g = gw + M::row(lk, r);
            double f, dx, dy, dz;
            gv.get((size_t)(g < G ? g : G - 1) * nbf + cc, f, dx, dy, dz);
            const double x = (cin && g < G) ? (double)acc[t][r] : 0.0;
            rs[r] += x * f;
            xs[r] += x * dx;
            ys[r] += x * dy;
            zs[r] += x * dz;
        }
        // one tile's loads in flight at a time: hoisting all of them costs hundreds of VGPRs
        asm volatile("" ::: "memory");
    }
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
    // per column half, row tile and point: rho, grad rho sums, in the D staging area (free now)
    static_assert(sizeof(Bs) >= 2 * 4 * kRhoPW * 4 * sizeof(double), "red2 fits in Bs");
    auto red2 = reinterpret_cast<double (*)[4][kRhoPW][4]>(&Bs[0][0]);
    __syncthreads();  // every wavefront's last fragment reads are done
    if (li == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int p = M::row(lk, r);
            red2[half][rt][p][0] = rs[r];
            red2[half][rt][p][1] = xs[r];
            red2[half][rt][p][2] = ys[r];
            red2[half][rt][p][3] = zs[r];
        }
    }
    __syncthreads();
    // the two column halves, in a fixed order
    if (tid < 4 * kRhoPW * 4) {
        const int q = tid & 3, p = (tid >> 2) % kRhoPW, r4 = tid / (4 * kRhoPW);
        red[r4][p][q] = red2[0][r4][p][q] + red2[1][r4][p][q];
    }
    __syncthreads();
    xc_point_tail<T, 8, (16 * NT + 63) / 64>(G, nbf, wv, l, g0, w, gv, aow, exc_sum, red, coef, part);
}

// Vparts[chunk] = Phi[chunk rows]^T aow[chunk rows] by 32 VW square output tiles (each wavefront a
// 16 VW square quarter: every fragment read feeds VW MFMAs); block -> (chunk, tile) so that the
// tiles of one chunk share an XCD (blocks are dealt round-robin over the 8 XCDs) and reuse its Phi
// and aow rows from that XCD's L2.
template <typename T, int VW>
__global__ __launch_bounds__(256) void xc_v_kernel(int G, int nbf, int ntile, int nchunk, int chunk,
                                                   const T* __restrict__ phi, const T* __restrict__ aow,
                                                   double* __restrict__ Vparts) {
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int kVW = VW, kVBM = 32 * VW;
    static_assert(kVBK * kVBM % 256 == 0, "whole elements per thread in the staging");
    __shared__ T As[kVBK][kVBM + kPad];  // Phi[g][m0 + i]
    __shared__ T Bs[kVBK][kVBM + kPad];  // aow[g][n0 + j]
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
    const int wr = wv >> 1, wc = wv & 1;  // this wavefront's quarter of the tile
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nt2 = ntile * ntile;
    const int ch = (slot / nt2) * 8 + xcd, tile = slot % nt2;
    if (ch >= nchunk) return;  // the whole block: before any barrier
    const int m0 = (tile / ntile) * kVBM, n0 = (tile % ntile) * kVBM;
    const int gb = ch * chunk, ge = min(G, gb + chunk);
    acc_t acc[kVW][kVW];
#pragma unroll
    for (int i = 0; i < kVW; ++i)
#pragma unroll
        for (int j = 0; j < kVW; ++j) acc[i][j] = acc_t{T(0), T(0), T(0), T(0)};
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
            T a[kVW], b[kVW];
#pragma unroll
            for (int i = 0; i < kVW; ++i) {
                a[i] = As[ks + lk][16 * kVW * wr + 16 * i + li];
                b[i] = Bs[ks + lk][16 * kVW * wc + 16 * i + li];
            }
            // 16 x 16 sub-tiles wholly past nbf (the edge tiles of the 128 x 128 grid) are skipped:
            // for nbf = 304 that is 37 % of the products; the condition is uniform per wavefront
#pragma unroll
            for (int i = 0; i < kVW; ++i)
#pragma unroll
                for (int j = 0; j < kVW; ++j)
                    if (m0 + 16 * kVW * wr + 16 * i < nbf && n0 + 16 * kVW * wc + 16 * j < nbf)
                        acc[i][j] = M::mma(a[i], b[j], acc[i][j]);
        }
    }
    double* out = Vparts + (size_t)ch * nbf * nbf;
#pragma unroll
    for (int i = 0; i < kVW; ++i)
#pragma unroll
        for (int j = 0; j < kVW; ++j) {
            const int col = n0 + 16 * kVW * wc + 16 * j + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + 16 * kVW * wr + 16 * i + M::row(lk, r);
                if (row < nbf && col < nbf) out[(size_t)row * nbf + col] = (double)acc[i][j][r];
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

__global__ void to_f32x4_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                const double* __restrict__ c, const double* __restrict__ d, float4* __restrict__ out,
                                size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4((float)a[i], (float)b[i], (float)c[i], (float)d[i]);
}

hipError_t launch_to_f32x4(const double* a, const double* b, const double* c, const double* d, float4* out, size_t n,
                           hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(to_f32x4_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, c, d, out, n);
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
                                   const T* dpy, const T* dpz, const float4* g4, const T* Dm, T* aow, double* exc_sum,
                                   double* Vparts, hipStream_t st) {
    const GridVals<T> gv{phi, dpx, dpy, dpz, g4};
    if (G <= 0 || nbf <= 0 || nchunk <= 0) return hipErrorInvalidValue;
    // column passes of at most 5 MFMA tiles (4 for FP64, whose accumulators take twice the
    // registers), with the tiles spread evenly over the passes.  FP32: 5 tiles keep the kernel at
    // 164 registers, three wavefronts per SIMD; 7 or 8 tiles need about 200 (two wavefronts) and
    // were 7 % slower on (H2O)16, 4 tiles 1.5 % slower.
    constexpr int kMaxNT = sizeof(T) == 8 ? 4 : 5;
    const int tiles = (nbf + 15) / 16, passes = (tiles + kMaxNT - 1) / kMaxNT, nt = (tiles + passes - 1) / passes;
    const dim3 grid((G + kRhoBM - 1) / kRhoBM);
    // up to kRhoAllMaxNT column tiles: the all-columns kernel (Phi read once per K step)
    switch (tiles <= kRhoAllMaxNT ? 100 + tiles : nt) {
#define QCG_RHO_ALL(n)                                                                                  \
    case 100 + n:                                                                                       \
        hipLaunchKernelGGL((xc_rho_all_kernel<T, n>), grid, dim3(512), 0, st, G, nbf, w, phi, gv, Dm, aow,  \
                           exc_sum);                                                                    \
        break;
        QCG_RHO_ALL(1) QCG_RHO_ALL(2) QCG_RHO_ALL(3) QCG_RHO_ALL(4) QCG_RHO_ALL(5) QCG_RHO_ALL(6) QCG_RHO_ALL(7)
        QCG_RHO_ALL(8) QCG_RHO_ALL(9) QCG_RHO_ALL(10) QCG_RHO_ALL(11) QCG_RHO_ALL(12) QCG_RHO_ALL(13)
        QCG_RHO_ALL(14) QCG_RHO_ALL(15) QCG_RHO_ALL(16) QCG_RHO_ALL(17) QCG_RHO_ALL(18) QCG_RHO_ALL(19)
        QCG_RHO_ALL(20) QCG_RHO_ALL(21) QCG_RHO_ALL(22)
#undef QCG_RHO_ALL
#define QCG_RHO_NT(n)                                                                                   \
    case n:                                                                                             \
        hipLaunchKernelGGL((xc_rho_kernel<T, n>), grid, dim3(256), 0, st, G, nbf, w, phi, gv, Dm, aow,      \
                           exc_sum);                                                                    \
        break;
        QCG_RHO_NT(1) QCG_RHO_NT(2) QCG_RHO_NT(3) QCG_RHO_NT(4) QCG_RHO_NT(5)
#undef QCG_RHO_NT
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int chunk = (G + nchunk - 1) / nchunk;
    const int nch = (G + chunk - 1) / chunk;  // <= nchunk partial matrices written
    const int vw = xc_v_width(nbf, sizeof(T) == 4);
    const int ntile = (nbf + 32 * vw - 1) / (32 * vw);
    const int nblk = 8 * ((nch + 7) / 8) * ntile * ntile;
    switch (vw) {
        case 3:
            hipLaunchKernelGGL((xc_v_kernel<T, 3>), dim3(nblk), dim3(256), 0, st, G, nbf, ntile, nch, chunk, phi, aow,
                               Vparts);
            break;
        case 4:
            hipLaunchKernelGGL((xc_v_kernel<T, 4>), dim3(nblk), dim3(256), 0, st, G, nbf, ntile, nch, chunk, phi, aow,
                               Vparts);
            break;
        case 5:
            hipLaunchKernelGGL((xc_v_kernel<T, 5>), dim3(nblk), dim3(256), 0, st, G, nbf, ntile, nch, chunk, phi, aow,
                               Vparts);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
static int v_blocks_per_cu_t(int vw) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (vw == 3) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, xc_v_kernel<T, 3>, 256, 0);
    else if (vw == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, xc_v_kernel<T, 4>, 256, 0);
    else if (vw == 5) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, xc_v_kernel<T, 5>, 256, 0);
    return e == hipSuccess && n > 0 ? n : 1;
}

int xc_v_blocks_per_cu(int nbf, bool fp32) {
    const int vw = xc_v_width(nbf, fp32);
    return fp32 ? v_blocks_per_cu_t<float>(vw) : v_blocks_per_cu_t<double>(vw);
}

hipError_t launch_xc_mfma(int G, int nbf, int nchunk, const double* w, const float* phi, const float4* g4,
                          const float* Dm, float* aow, double* exc_sum, double* Vparts, hipStream_t st) {
    return launch_xc_mfma_t<float>(G, nbf, nchunk, w, phi, nullptr, nullptr, nullptr, g4, Dm, aow, exc_sum, Vparts,
                                   st);
}

hipError_t launch_xc_mfma(int G, int nbf, int nchunk, const double* w, const double* phi, const double* dpx,
                          const double* dpy, const double* dpz, const double* Dm, double* aow, double* exc_sum,
                          double* Vparts, hipStream_t st) {
    return launch_xc_mfma_t<double>(G, nbf, nchunk, w, phi, dpx, dpy, dpz, nullptr, Dm, aow, exc_sum, Vparts, st);
}

}  // namespace qcg
