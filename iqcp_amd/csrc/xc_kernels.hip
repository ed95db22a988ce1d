// Exchange-correlation kernels on the grid (B3LYP, closed shell):
//   basis_kernel   : Phi and grad Phi on the grid once per SCF (evaluate_basis_and_gradients_on_grid,
//                    dft/ks_scf.rs:115-255): per-primitive Gaussians, per-component normalisation,
//                    screening alpha_min r^2 > 34.5388 (orbital/grid.rs:174), |angular| < 1e-30 -> 0.
//   point_kernel   : one wavefront per grid point, after X = Phi D (GEMM):
//                    rho = max(0, sum_mu Phi X), grad rho = 2 sum_mu grad Phi X, sigma = |grad rho|^2
//                    (compute_density_and_gradient_matmul, :1028-1093), B3LYP exc/vrho/vsigma, and the
//                    fused aow = w vrho Phi + 4 w vsigma (grad rho . grad Phi) row (build_vxc_gga_matmul,
//                    :1150-1296), plus E_xc = sum w exc rho.
// The two GEMMs (X = Phi D, V = Phi^T aow) run between them; V is symmetrised at the end.
#include "xc_functional.h"
#include "xc_launch.h"

namespace qcg {

__global__ __launch_bounds__(256) void basis_kernel(int G, int nshell, const GridShell* __restrict__ shells,
                                                    const double* __restrict__ exps, const double* __restrict__ cn,
                                                    const double* __restrict__ pts, int nbf, double* __restrict__ phi,
                                                    double* __restrict__ dphx, double* __restrict__ dphy,
                                                    double* __restrict__ dphz) {
    const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (idx >= (size_t)G * nshell) return;
    const int g = (int)(idx / nshell), s = (int)(idx % nshell);
    const GridShell sh = shells[s];
    const double dx = pts[3 * (size_t)g] - sh.cx, dy = pts[3 * (size_t)g + 1] - sh.cy, dz = pts[3 * (size_t)g + 2] - sh.cz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const int nc = ncart(sh.l);
    const size_t row = (size_t)g * nbf + sh.bf;
    if (sh.amin * r2 > 34.5388) {  // GAUSSIAN_SCREENING_THRESHOLD (orbital/grid.rs:174)
        for (int c = 0; c < nc; ++c) phi[row + c] = dphx[row + c] = dphy[row + c] = dphz[row + c] = 0.0;
        return;
    }
    // radial sums: R0 = sum c_p e_p, R1 = sum c_p alpha_p e_p
    double R0 = 0.0, R1 = 0.0;
    for (int p = 0; p < sh.nprim; ++p) {
        const double a = exps[sh.poff + p];
        const double e = cn[sh.poff + p] * exp(-a * r2);
        R0 += e;
        R1 += a * e;
    }
    const double d[3] = {dx, dy, dz};
    for (int c = 0; c < nc; ++c) {
        const int px = cart_x(sh.l, c), py = cart_y(sh.l, c), pz = cart_z(sh.l, c);
        const double nrm = (sh.l == 2 && (px == 2 || py == 2 || pz == 2)) ? 0.57735026918962576451 : 1.0;
        auto pw = [](double v, int k) { return k == 0 ? 1.0 : (k == 1 ? v : v * v); };
        const double ang = pw(dx, px) * pw(dy, py) * pw(dz, pz);
        phi[row + c] = fabs(ang) >= 1e-30 ? nrm * ang * R0 : 0.0;
        const int pp[3] = {px, py, pz};
        double gr[3];
        for (int k = 0; k < 3; ++k) {
            double dang = 0.0;
            if (pp[k] > 0) {
                int q[3] = {px, py, pz};
                q[k] -= 1;
                dang = pp[k] * pw(dx, q[0]) * pw(dy, q[1]) * pw(dz, q[2]);
            }
            gr[k] = nrm * (dang * R0 - 2.0 * d[k] * ang * R1);
        }
        dphx[row + c] = gr[0];
        dphy[row + c] = gr[1];
        dphz[row + c] = gr[2];
    }
}

__global__ __launch_bounds__(256) void point_kernel(int G, int nbf, const double* __restrict__ w,
                                                    const double* __restrict__ phi, const double* __restrict__ dphx,
                                                    const double* __restrict__ dphy, const double* __restrict__ dphz,
                                                    const double* __restrict__ X, double* __restrict__ aow,
                                                    double* __restrict__ exc_sum, double* __restrict__ rho_out) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    __shared__ double part[4];
    double contrib = 0.0;
    if (g < G) {
        const size_t row = (size_t)g * nbf;
        double r = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
        for (int mu = lane; mu < nbf; mu += 64) {
            const double x = X[row + mu];
            r += phi[row + mu] * x;
            gx += dphx[row + mu] * x;
            gy += dphy[row + mu] * x;
            gz += dphz[row + mu] * x;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            r += __shfl_xor(r, off, 64);
            gx += __shfl_xor(gx, off, 64);
            gy += __shfl_xor(gy, off, 64);
            gz += __shfl_xor(gz, off, 64);
        }
        const double rho = fmax(r, 0.0);
        const double grx = 2.0 * gx, gry = 2.0 * gy, grz = 2.0 * gz;
        const double sigma = grx * grx + gry * gry + grz * grz;
        double exc, vrho, vsig;
        xc_b3lyp(rho, sigma, exc, vrho, vsig);
        const double wg = w[g];
        const double wvr = wg * vrho, f4 = 4.0 * wg * vsig;
        for (int mu = lane; mu < nbf; mu += 64)
            aow[row + mu] = wvr * phi[row + mu] + f4 * (grx * dphx[row + mu] + gry * dphy[row + mu] + grz * dphz[row + mu]);
        contrib = wg * exc * rho;
        if (lane == 0 && rho_out) rho_out[g] = rho;
    }
    if (lane == 0) part[threadIdx.x >> 6] = contrib;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += part[i];
        if (s != 0.0) atomicAdd(exc_sum, s);
    }
}

__global__ void xc_eval_kernel(int n, const double* __restrict__ rho, const double* __restrict__ sigma,
                               double* __restrict__ exc, double* __restrict__ vrho, double* __restrict__ vsigma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    xc_b3lyp(rho[i], sigma[i], exc[i], vrho[i], vsigma[i]);
}

__global__ void sum_parts_kernel(const double* __restrict__ parts, int nparts, size_t n, double* __restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < nparts; ++k) s += parts[(size_t)k * n + i];
    out[i] = s;
}

hipError_t launch_sum_parts(const double* parts, int nparts, size_t n, double* out, hipStream_t st) {
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, parts, nparts, n, out);
    return hipGetLastError();
}

__global__ void sym_half_kernel(const double* __restrict__ V, double* __restrict__ out, int n) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    out[(size_t)i * n + j] = 0.5 * (V[(size_t)i * n + j] + V[(size_t)j * n + i]);
}

// XC nuclear gradient (xc_gradient_analytical, scf/gradient.rs:1687-1882; B3LYP, no grid-weight
// response).  Per point: rho, grad rho, B3LYP, then the row
//   A_g = w vrho Phi_g + 2 w vsigma (grad rho . grad Phi_g)
// so that (A D)_g,mu = w vrho (chi D)_mu + 2 w vsigma grad rho . (grad chi D)_mu is the reference's
// aow (:1846-1856), and t_g = 2 w vsigma grad rho for the basis-Hessian part (:1858-1874).
__global__ __launch_bounds__(256) void xc_grad_point_kernel(int G, int nbf, const double* __restrict__ w,
                                                            const double* __restrict__ phi,
                                                            const double* __restrict__ dphx,
                                                            const double* __restrict__ dphy,
                                                            const double* __restrict__ dphz,
                                                            const double* __restrict__ X, double* __restrict__ A,
                                                            double* __restrict__ t) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (g >= G) return;
    const size_t row = (size_t)g * nbf;
    double r = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int mu = lane; mu < nbf; mu += 64) {
        const double x = X[row + mu];
        r += phi[row + mu] * x;
        gx += dphx[row + mu] * x;
        gy += dphy[row + mu] * x;
        gz += dphz[row + mu] * x;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        r += __shfl_xor(r, off, 64);
        gx += __shfl_xor(gx, off, 64);
        gy += __shfl_xor(gy, off, 64);
        gz += __shfl_xor(gz, off, 64);
    }
    const double rho = fmax(r, 0.0);
    const double grx = 2.0 * gx, gry = 2.0 * gy, grz = 2.0 * gz;
    double exc, vrho, vsig;
    xc_b3lyp(rho, grx * grx + gry * gry + grz * grz, exc, vrho, vsig);
    const double wvr = w[g] * vrho, f2 = 2.0 * w[g] * vsig;
    for (int mu = lane; mu < nbf; mu += 64)
        A[row + mu] = wvr * phi[row + mu] + f2 * (grx * dphx[row + mu] + gry * dphy[row + mu] + grz * dphz[row + mu]);
    if (lane == 0) {
        t[3 * (size_t)g] = f2 * grx;
        t[3 * (size_t)g + 1] = f2 * gry;
        t[3 * (size_t)g + 2] = f2 * grz;
    }
}

// One block row per shell (blockIdx.y), threads over grid points: every thread of a block adds to
// the same atom, so the three sums are wave-reduced before one atomic per wave.
//   grad[A] += -2 (A D)_g,mu grad chi_mu(g) - 2 (chi D)_g,mu sum_e t_e d2 chi_mu/dr_d dr_e
// with d2 chi/dr_a dr_b = N [d2ang R0 - 2 (r_b dang_a + r_a dang_b + delta_ab ang) R1
// + 4 r_a r_b ang R2], R_k = sum_p c_p n_p alpha_p^k e^{-alpha_p r^2} (gradient.rs:1938-2050).
__global__ __launch_bounds__(256) void xc_grad_kernel(int G, const GridShell* __restrict__ shells,
                                                      const double* __restrict__ exps, const double* __restrict__ cn,
                                                      const double* __restrict__ pts, int nbf,
                                                      const double* __restrict__ X, const double* __restrict__ M,
                                                      const double* __restrict__ t, double* __restrict__ grad) {
    const GridShell sh = shells[blockIdx.y];
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (g < G) {
        const double d[3] = {pts[3 * (size_t)g] - sh.cx, pts[3 * (size_t)g + 1] - sh.cy, pts[3 * (size_t)g + 2] - sh.cz};
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(sh.amin * r2 > 34.5388)) {
            double R0 = 0.0, R1 = 0.0, R2 = 0.0;
            for (int p = 0; p < sh.nprim; ++p) {
                const double a = exps[sh.poff + p];
                const double e = cn[sh.poff + p] * exp(-a * r2);
                R0 += e;
                R1 += a * e;
                R2 += a * a * e;
            }
            const double tv[3] = {t[3 * (size_t)g], t[3 * (size_t)g + 1], t[3 * (size_t)g + 2]};
            const size_t row = (size_t)g * nbf + sh.bf;
            auto pw = [](double v, int k) { return k <= 0 ? (k == 0 ? 1.0 : 0.0) : (k == 1 ? v : v * v); };
            const int nc = ncart(sh.l);
            for (int c = 0; c < nc; ++c) {
                const int p3[3] = {cart_x(sh.l, c), cart_y(sh.l, c), cart_z(sh.l, c)};
                const double nrm = (sh.l == 2 && (p3[0] == 2 || p3[1] == 2 || p3[2] == 2)) ? 0.57735026918962576451 : 1.0;
                auto angp = [&](int a0, int a1, int a2) { return pw(d[0], a0) * pw(d[1], a1) * pw(d[2], a2); };
                const double ang = angp(p3[0], p3[1], p3[2]);
                double dang[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    int q[3] = {p3[0], p3[1], p3[2]};
                    q[a] -= 1;
                    dang[a] = p3[a] > 0 ? p3[a] * angp(q[0], q[1], q[2]) : 0.0;
                }
                const double xv = X[row + c], mv = M[row + c];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double gphi = nrm * (dang[a] * R0 - 2.0 * d[a] * ang * R1);
                    double hc = 0.0;
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        double d2 = 0.0;
                        int q[3] = {p3[0], p3[1], p3[2]};
                        if (a == b) {
                            if (p3[a] >= 2) { q[a] -= 2; d2 = p3[a] * (p3[a] - 1) * angp(q[0], q[1], q[2]); }
                        } else if (p3[a] > 0 && p3[b] > 0) {
                            q[a] -= 1; q[b] -= 1;
                            d2 = p3[a] * p3[b] * angp(q[0], q[1], q[2]);
                        }
                        const double h = nrm * (d2 * R0 - 2.0 * (d[b] * dang[a] + d[a] * dang[b] + (a == b ? ang : 0.0)) * R1 +
                                                4.0 * d[a] * d[b] * ang * R2);
                        hc += tv[b] * h;
                    }
                    acc[a] += -2.0 * mv * gphi - 2.0 * xv * hc;
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v = acc[a];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(&grad[3 * sh.atom + a], v);
    }
}

hipError_t launch_xc_grad_points(int G, int nbf, const double* w, const double* phi, const double* dphx,
                                 const double* dphy, const double* dphz, const double* X, double* A, double* t,
                                 hipStream_t st) {
    if (G == 0) return hipSuccess;
    hipLaunchKernelGGL(xc_grad_point_kernel, dim3((G + 3) / 4), dim3(256), 0, st, G, nbf, w, phi, dphx, dphy, dphz, X,
                       A, t);
    return hipGetLastError();
}

hipError_t launch_xc_grad_accumulate(int G, int nshell, const GridShell* shells, const double* exps, const double* cn,
                                     const double* pts, int nbf, const double* X, const double* M, const double* t,
                                     double* grad, hipStream_t st) {
    if (G == 0 || nshell == 0) return hipSuccess;
    hipLaunchKernelGGL(xc_grad_kernel, dim3((G + 255) / 256, nshell), dim3(256), 0, st, G, shells, exps, cn, pts, nbf,
                       X, M, t, grad);
    return hipGetLastError();
}

// Becke partition weight of one grid point per thread (becke_partition, dft/grid.rs:400-468): the
// same operations in the same order as the host loop in grid.cpp, with contraction into FMA off on
// both sides, so the kept-point decision |w| > 1e-30 matches.  Each thread keeps its nuclear
// distances and cell products in LDS (2 nat doubles).
__global__ __launch_bounds__(64) void becke_partition_kernel(int64_t np, const double* __restrict__ pts,
                                                             const int* __restrict__ owner, int nat,
                                                             const double* __restrict__ axyz,
                                                             const double* __restrict__ dist, double* __restrict__ part) {
#pragma clang fp contract(off)
    extern __shared__ double sh[];
    double* gd = sh + (size_t)threadIdx.x * 2 * nat;
    double* P = gd + nat;
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= np) return;  // no barrier below
    for (int i = 0; i < nat; ++i) {
        const double dx = pts[3 * p] - axyz[3 * i], dy = pts[3 * p + 1] - axyz[3 * i + 1],
                     dz = pts[3 * p + 2] - axyz[3 * i + 2];
        gd[i] = sqrt(dx * dx + dy * dy + dz * dz);
        P[i] = 1.0;
    }
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j < i; ++j) {
            const double rij = dist[(size_t)i * nat + j];
            if (rij < 1e-15) continue;
            double mu = (gd[i] - gd[j]) / rij;
            for (int it = 0; it < 3; ++it) mu = mu * (3.0 - mu * mu) * 0.5;
            P[i] *= 0.5 * (1.0 - mu);
            P[j] *= 0.5 * (1.0 + mu);
        }
    double ps = 0.0;
    for (int i = 0; i < nat; ++i) ps += P[i];
    part[p] = fabs(ps) > 1e-30 ? P[owner[p]] / ps : 1.0 / nat;
}

hipError_t launch_becke_partition(int64_t np, const double* pts, const int* owner, int nat, const double* axyz,
                                  const double* dist, double* part, hipStream_t st) {
    if (np <= 0) return hipSuccess;
    if (nat <= 0 || nat > 4096) return hipErrorInvalidValue;
    int tpb = 64;  // threads per block: 2 nat doubles each, at most 64 KB of LDS per block
    while (tpb > 1 && (size_t)tpb * 2 * nat * sizeof(double) > 65536) tpb >>= 1;
    const size_t lds = (size_t)tpb * 2 * nat * sizeof(double);
    hipLaunchKernelGGL(becke_partition_kernel, dim3((unsigned)((np + tpb - 1) / tpb)), dim3(tpb), lds, st, np, pts,
                       owner, nat, axyz, dist, part);
    return hipGetLastError();
}

hipError_t launch_basis_on_grid(int G, int nshell, const GridShell* shells, const double* exps, const double* cn,
                                const double* pts, int nbf, double* phi, double* dphx, double* dphy, double* dphz,
                                hipStream_t st) {
    const size_t n = (size_t)G * nshell;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(basis_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, G, nshell, shells, exps, cn,
                       pts, nbf, phi, dphx, dphy, dphz);
    return hipGetLastError();
}

hipError_t launch_xc_points(int G, int nbf, const double* w, const double* phi, const double* dphx,
                            const double* dphy, const double* dphz, const double* X, double* aow, double* exc_sum,
                            double* rho_out, hipStream_t st) {
    if (G == 0) return hipSuccess;
    hipLaunchKernelGGL(point_kernel, dim3((G + 3) / 4), dim3(256), 0, st, G, nbf, w, phi, dphx, dphy, dphz, X, aow,
                       exc_sum, rho_out);
    return hipGetLastError();
}

hipError_t launch_xc_eval(int n, const double* rho, const double* sigma, double* exc, double* vrho, double* vsigma,
                          hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(xc_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, rho, sigma, exc, vrho, vsigma);
    return hipGetLastError();
}

hipError_t launch_sym_half(const double* V, double* out, int n, hipStream_t st) {
    hipLaunchKernelGGL(sym_half_kernel, dim3((n + 127) / 128, n), dim3(128), 0, st, V, out, n);
    return hipGetLastError();
}

}  // namespace qcg
