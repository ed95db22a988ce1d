// Host-side launch interface of the XC kernels.
#pragma once
#include "qcg_common.h"

namespace qcg {

struct GridShell {
    double cx, cy, cz, amin;  // centre, smallest exponent (screening)
    int l, nprim, poff, bf;
    int atom, pad;
};

hipError_t launch_basis_on_grid(int G, int nshell, const GridShell* shells, const double* exps, const double* cn,
                                const double* pts, int nbf, double* phi, double* dphx, double* dphy, double* dphz,
                                hipStream_t st);
hipError_t launch_xc_points(int G, int nbf, const double* w, const double* phi, const double* dphx,
                            const double* dphy, const double* dphz, const double* X, double* aow, double* exc_sum,
                            double* rho_out, hipStream_t st);
hipError_t launch_xc_eval(int n, const double* rho, const double* sigma, double* exc, double* vrho, double* vsigma,
                          hipStream_t st);
hipError_t launch_sym_half(const double* V, double* out, int n, hipStream_t st);
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
