// One-electron integrals on the host (setup work, FP64): overlap S, kinetic T, nuclear attraction V,
// and their nuclear derivatives for the gradient.
//
// Same matrices as qc-core's overlap_matrix (integrals/overlap.rs:647), kinetic (integrals/kinetic.rs),
// nuclear attraction (integrals/nuclear.rs:226, Rys quadrature) and hcore_matrix (integrals/hcore.rs:103),
// with the reference's per-primitive, per-Cartesian-component normalisation
// (cartesian_gaussian_normalization, integrals/overlap.rs:343-377) and raw contraction coefficients.
// Overlap/kinetic use Obara-Saika 1-D recurrences; V uses Rys quadrature with the same node table
// as the two-electron kernels.
#include "onee.h"

#include <cmath>
#include <vector>

#include "rys.h"

namespace qcg {

namespace {

double prim_norm_radial(double a, int l) {
    const double two_a_pi = 2.0 * a / kPi;
    const double base = std::sqrt(std::sqrt(two_a_pi * two_a_pi * two_a_pi));
    const double fa = 4.0 * a;
    double num = 1.0;
    if (l == 1) num = std::sqrt(fa);
    else if (l == 2) num = fa;
    else if (l > 2) num = std::pow(fa, l / 2.0);
    return base * num;
}

double comp_norm_host(int l, int c) {
    const int x = cart_x(l, c), y = cart_y(l, c), z = cart_z(l, c);
    return 1.0 / std::sqrt(dfact_odd(x) * dfact_odd(y) * dfact_odd(z));
}

// 1-D overlap table s[i][j], i <= imax, j <= jmax (Obara-Saika), without the exp/sqrt prefactor.
void os_overlap_1d(double PA, double PB, double oo2p, int imax, int jmax, double* s, int ld) {
    s[0] = 1.0;
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j) {
            if (i == 0 && j == 0) continue;
            double v;
            if (i > 0) {
                v = PA * s[(i - 1) * ld + j];
                if (i > 1) v += (i - 1) * oo2p * s[(i - 2) * ld + j];
                if (j > 0) v += j * oo2p * s[(i - 1) * ld + j - 1];
            } else {
                v = PB * s[i * ld + j - 1];
                if (j > 1) v += (j - 1) * oo2p * s[i * ld + j - 2];
            }
            s[i * ld + j] = v;
        }
}

}  // namespace

// Shell-pair blocks are written into full nbf x nbf row-major matrices.
void one_electron_matrices(const HostBasis& B, double* S, double* T, double* V) {
    const int nbf = B.nbf;
    for (size_t i = 0; i < (size_t)nbf * nbf; ++i) {
        if (S) S[i] = 0.0;
        if (T) T[i] = 0.0;
        if (V) V[i] = 0.0;
    }
    const double* rys_tab = qcg_rys_coef;
    for (int si = 0; si < B.nshell; ++si)
        for (int sj = 0; sj < B.nshell; ++sj) {
            const int la = B.l[si], lb = B.l[sj];
            const int na = ncart(la), nb = ncart(lb);
            const double* A = &B.centers[3 * si];
            const double* Bc = &B.centers[3 * sj];
            const double AB[3] = {A[0] - Bc[0], A[1] - Bc[1], A[2] - Bc[2]};
            const double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
            for (int pi = 0; pi < B.nprim[si]; ++pi)
                for (int pj = 0; pj < B.nprim[sj]; ++pj) {
                    const double a = B.exps[B.prim_off[si] + pi], b = B.exps[B.prim_off[sj] + pj];
                    const double ca = B.coefs[B.prim_off[si] + pi] * prim_norm_radial(a, la);
                    const double cb = B.coefs[B.prim_off[sj] + pj] * prim_norm_radial(b, lb);
                    const double p = a + b, mu = a * b / p, oo2p = 0.5 / p;
                    const double kab = std::exp(-mu * ab2);
                    double P[3], PA[3], PB[3];
                    for (int x = 0; x < 3; ++x) {
                        P[x] = (a * A[x] + b * Bc[x]) / p;
                        PA[x] = P[x] - A[x];
                        PB[x] = P[x] - Bc[x];
                    }
                    // overlap tables up to (la+1, lb+1) for the kinetic terms
                    constexpr int LD = kMaxL + 4;
                    double s1[3][LD * LD];
                    for (int x = 0; x < 3; ++x) os_overlap_1d(PA[x], PB[x], oo2p, la + 1, lb + 1, s1[x], LD);
                    const double pre_s = kab * std::pow(kPi / p, 1.5) * ca * cb;
                    // nuclear attraction via Rys: n = (la+lb)/2 + 1 nodes per nucleus
                    const int nr = (la + lb) / 2 + 1;
                    for (int ia = 0; ia < na; ++ia)
                        for (int jb = 0; jb < nb; ++jb) {
                            const int ix = cart_x(la, ia), iy = cart_y(la, ia), iz = cart_z(la, ia);
                            const int jx = cart_x(lb, jb), jy = cart_y(lb, jb), jz = cart_z(lb, jb);
                            const double nrm = comp_norm_host(la, ia) * comp_norm_host(lb, jb);
                            const size_t idx = (size_t)(B.bf_off[si] + ia) * nbf + B.bf_off[sj] + jb;
                            const double sx = s1[0][ix * LD + jx], sy = s1[1][iy * LD + jy], sz = s1[2][iz * LD + jz];
                            if (S) S[idx] += pre_s * nrm * sx * sy * sz;
                            if (T) {
                                auto t1 = [&](int x, int i, int j) {
                                    const double* s = s1[x];
                                    double v = 4.0 * a * b * s[(i + 1) * LD + j + 1];
                                    if (i > 0) v -= 2.0 * b * i * s[(i - 1) * LD + j + 1];
                                    if (j > 0) v -= 2.0 * a * j * s[(i + 1) * LD + j - 1];
                                    if (i > 0 && j > 0) v += (double)(i * j) * s[(i - 1) * LD + j - 1];
                                    return 0.5 * v;
                                };
                                const double tx = t1(0, ix, jx), ty = t1(1, iy, jy), tz = t1(2, iz, jz);
                                T[idx] += pre_s * nrm * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
                            }
                        }
                    if (V) {
                        for (int c = 0; c < B.natom; ++c) {
                            const double* C = &B.xyz[3 * c];
                            const double PC[3] = {P[0] - C[0], P[1] - C[1], P[2] - C[2]};
                            const double Tn = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                            double rx[QCG_RYS_NMAX], rw[QCG_RYS_NMAX];
                            rys_eval_n<double>(nr, Tn, rys_tab, rx, rw);
                            const double pre_v = -(double)B.Z[c] * 2.0 * kPi / p * kab * ca * cb;
                            for (int r = 0; r < nr; ++r) {
                                const double u = rx[r];
                                // 1-D tables g[i][j] (i <= la, j <= lb) per axis
                                double g[3][(kMaxL + 1) * (kMaxL + 1)];
                                for (int x = 0; x < 3; ++x) {
                                    double h[2 * kMaxL + 2];
                                    const double c0 = PA[x] - u * PC[x];
                                    const double bb = (1.0 - u) * oo2p;
                                    h[0] = 1.0;
                                    if (la + lb > 0) h[1] = c0;
                                    for (int n = 1; n < la + lb; ++n) h[n + 1] = c0 * h[n] + n * bb * h[n - 1];
                                    double t2[(2 * kMaxL + 1) * (kMaxL + 1)];
                                    for (int n = 0; n <= la + lb; ++n) t2[n * (kMaxL + 1)] = h[n];
                                    for (int j = 1; j <= lb; ++j)
                                        for (int n = 0; n + j <= la + lb; ++n)
                                            t2[n * (kMaxL + 1) + j] = t2[(n + 1) * (kMaxL + 1) + j - 1] + AB[x] * t2[n * (kMaxL + 1) + j - 1];
                                    for (int i = 0; i <= la; ++i)
                                        for (int j = 0; j <= lb; ++j) g[x][i * (kMaxL + 1) + j] = t2[i * (kMaxL + 1) + j];
                                }
                                const double wr = rw[r] * pre_v;
                                for (int ia = 0; ia < na; ++ia)
                                    for (int jb = 0; jb < nb; ++jb) {
                                        const double nrm = comp_norm_host(la, ia) * comp_norm_host(lb, jb);
                                        const size_t idx = (size_t)(B.bf_off[si] + ia) * nbf + B.bf_off[sj] + jb;
                                        V[idx] += wr * nrm * g[0][cart_x(la, ia) * (kMaxL + 1) + cart_x(lb, jb)] *
                                                  g[1][cart_y(la, ia) * (kMaxL + 1) + cart_y(lb, jb)] *
                                                  g[2][cart_z(la, ia) * (kMaxL + 1) + cart_z(lb, jb)];
                                    }
                            }
                        }
                    }
                }
        }
}

}  // namespace qcg

namespace qcg {

// nuclear_repulsion_gradient (scf/gradient.rs:135-176): dE/dR_A = -sum_B Z_A Z_B (R_A - R_B)/|R_A - R_B|^3
void nuclear_repulsion_gradient(const HostBasis& B, double* grad) {
    for (int a = 0; a < B.natom; ++a)
        for (int b = 0; b < B.natom; ++b) {
            if (a == b) continue;
            double d[3], r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                d[x] = B.xyz[3 * a + x] - B.xyz[3 * b + x];
                r2 += d[x] * d[x];
            }
            const double r = std::sqrt(r2);
            if (r < 1e-10) continue;
            for (int x = 0; x < 3; ++x) grad[3 * a + x] -= (double)B.Z[a] * B.Z[b] * d[x] / (r2 * r);
        }
}

// sum_ab D_ab dH_ab/dR - W_ab dS_ab/dR, H = T + V, over all ordered shell pairs and their primitive
// pairs (onee_grad_overlap_kinetic and onee_grad_nuclear, shared with the device kernel).
void one_electron_gradient(const HostBasis& B, const double* D, const double* W, double* grad_out) {
    const BasisView V = basis_view(B);
    const int n3 = 3 * B.natom;
    // shell pairs spread over the host threads; each thread sums into its own gradient copy
#pragma omp parallel
    {
    std::vector<double> gloc(n3, 0.0);
    double* grad = gloc.data();
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (int si = 0; si < B.nshell; ++si)
        for (int sj = 0; sj < B.nshell; ++sj)
            for (int pi = 0; pi < B.nprim[si]; ++pi)
                for (int pj = 0; pj < B.nprim[sj]; ++pj) {
                    auto add = [&](int atom, int x, double v) { grad[3 * atom + x] += v; };
                    onee_grad_overlap_kinetic(V, si, sj, pi, pj, D, W, add);
                    for (int cn = 0; cn < B.natom; ++cn) onee_grad_nuclear(V, si, sj, pi, pj, cn, D, qcg_rys_coef, add);
                }
#pragma omp critical
    for (int t = 0; t < n3; ++t) grad_out[t] += gloc[t];
    }
}

}  // namespace qcg
