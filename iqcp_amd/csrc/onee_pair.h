// One-electron gradient terms of one primitive pair, shared by the host loop (onee.cpp) and the
// device kernel (onee_kernels.hip): the same operations in the same order on both sides.
//
// sum_ab D_ab dH_ab/dR - W_ab dS_ab/dR, H = T + V, for the primitive pair (pi, pj) of the ordered
// shell pair (si, sj) (overlap_derivative_matrix, scf/gradient.rs:188, and hcore_derivative_matrix
// :355, contracted as in rhf_gradient :930-950).  Basis-function derivatives:
// d/dA_x phi_i = 2 alpha phi_(i+1) - i phi_(i-1); for V the operator centre C gets -(dA + dB) by
// translational invariance of each three-centre integral.
#pragma once
#include <cmath>

#include "qcg_common.h"
#include "rys.h"

namespace qcg {

// Plain-pointer view of a basis (host vectors or device copies).
struct BasisView {
    int nshell, nbf, natom;
    const int *l, *nprim, *prim_off, *shell_atom, *bf_off, *Z;
    const double *exps, *coefs, *centers, *xyz;
};

// Radial normalisation (2a/pi)^(3/4) (4a)^(l/2) of one primitive (cartesian_gaussian_normalization
// without the component factor, integrals/overlap.rs:343-377).
QCG_HD double onee_norm_radial(double a, int l) {
    const double two_a_pi = 2.0 * a / kPi;
    const double base = sqrt(sqrt(two_a_pi * two_a_pi * two_a_pi));
    const double fa = 4.0 * a;
    double num = 1.0;
    if (l == 1) num = sqrt(fa);
    else if (l == 2) num = fa;
    else if (l > 2) num = pow(fa, l / 2.0);
    return base * num;
}

QCG_HD double onee_norm_comp(int l, int c) {
    return 1.0 / sqrt(dfact_odd(cart_x(l, c)) * dfact_odd(cart_y(l, c)) * dfact_odd(cart_z(l, c)));
}

// 1-D overlap table s[i][j], i <= imax, j <= jmax (Obara-Saika), without the exp/sqrt prefactor.
QCG_HD void onee_overlap_1d(double PA, double PB, double oo2p, int imax, int jmax, double* s, int ld) {
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

// Rys rule of n <= 3 nodes (the most the one-electron terms of s, p, d shells need, derivatives
// included): only these orders are instantiated, which keeps the kernels' registers down.
QCG_HD void onee_rys(int n, double T, const double* rys_tab, double* x, double* w) {
    static_assert(kMaxL <= 2, "one-electron Rys orders above 3");
    if (n == 1) rys_eval<double, 1>(T, rys_tab, x, w);
    else if (n == 2) rys_eval<double, 2>(T, rys_tab, x, w);
    else rys_eval<double, 3>(T, rys_tab, x, w);
}

// Geometry and contraction data of one primitive pair.
struct OneePair {
    int la, lb, na, nb, atA, atB, bfa, bfb;
    double a, b, p, oo2p, kab, ca, cb;
    double AB[3], P[3], PA[3], PB[3];
};

QCG_HD OneePair onee_pair_setup(const BasisView& B, int si, int sj, int pi, int pj) {
    OneePair q;
    q.la = B.l[si];
    q.lb = B.l[sj];
    q.na = ncart(q.la);
    q.nb = ncart(q.lb);
    q.atA = B.shell_atom[si];
    q.atB = B.shell_atom[sj];
    q.bfa = B.bf_off[si];
    q.bfb = B.bf_off[sj];
    const double* A = &B.centers[3 * si];
    const double* Bc = &B.centers[3 * sj];
    for (int x = 0; x < 3; ++x) q.AB[x] = A[x] - Bc[x];
    const double ab2 = q.AB[0] * q.AB[0] + q.AB[1] * q.AB[1] + q.AB[2] * q.AB[2];
    q.a = B.exps[B.prim_off[si] + pi];
    q.b = B.exps[B.prim_off[sj] + pj];
    q.ca = B.coefs[B.prim_off[si] + pi] * onee_norm_radial(q.a, q.la);
    q.cb = B.coefs[B.prim_off[sj] + pj] * onee_norm_radial(q.b, q.lb);
    q.p = q.a + q.b;
    q.oo2p = 0.5 / q.p;
    q.kab = exp(-(q.a * q.b / q.p) * ab2);
    for (int x = 0; x < 3; ++x) {
        q.P[x] = (q.a * A[x] + q.b * Bc[x]) / q.p;
        q.PA[x] = q.P[x] - A[x];
        q.PB[x] = q.P[x] - Bc[x];
    }
    return q;
}

// Overlap and kinetic terms of the primitive pair (pi, pj) of the ordered shell pair (si, sj);
// add(atom, axis, value) accumulates a gradient contribution.
template <typename Add>
QCG_HD void onee_grad_overlap_kinetic(const BasisView& B, int si, int sj, int pi, int pj, const double* D,
                                      const double* W, Add&& add) {
    constexpr int LD = kMaxL + 4;
    const OneePair q = onee_pair_setup(B, si, sj, pi, pj);
    const double a = q.a, b = q.b;
    double s1[3][LD * LD];
    for (int x = 0; x < 3; ++x) onee_overlap_1d(q.PA[x], q.PB[x], q.oo2p, q.la + 2, q.lb + 2, s1[x], LD);
    auto sv = [&](int x, int i, int j) { return (i < 0 || j < 0) ? 0.0 : s1[x][i * LD + j]; };
    auto tv = [&](int x, int i, int j) {
        if (i < 0 || j < 0) return 0.0;
        double v = 4.0 * a * b * sv(x, i + 1, j + 1);
        if (i > 0) v -= 2.0 * b * i * sv(x, i - 1, j + 1);
        if (j > 0) v -= 2.0 * a * j * sv(x, i + 1, j - 1);
        if (i > 0 && j > 0) v += (double)(i * j) * sv(x, i - 1, j - 1);
        return 0.5 * v;
    };
    const double pre = q.kab * pow(kPi / q.p, 1.5) * q.ca * q.cb;
    double gA0[3] = {0, 0, 0}, gB0[3] = {0, 0, 0};
    for (int ia = 0; ia < q.na; ++ia)
        for (int jb = 0; jb < q.nb; ++jb) {
            const int ip[3] = {cart_x(q.la, ia), cart_y(q.la, ia), cart_z(q.la, ia)};
            const int jp[3] = {cart_x(q.lb, jb), cart_y(q.lb, jb), cart_z(q.lb, jb)};
            const size_t idx = (size_t)(q.bfa + ia) * B.nbf + q.bfb + jb;
            const double dv = D[idx], wv = W[idx];
            if (dv == 0.0 && wv == 0.0) continue;
            const double nrm = onee_norm_comp(q.la, ia) * onee_norm_comp(q.lb, jb);
            for (int x = 0; x < 3; ++x) {
                const int y = (x + 1) % 3, z = (x + 2) % 3;
                const double sy = sv(y, ip[y], jp[y]), sz = sv(z, ip[z], jp[z]);
                const double ty = tv(y, ip[y], jp[y]), tz = tv(z, ip[z], jp[z]);
                // derivative with respect to A_x and B_x of S and T
                const double dsA = 2.0 * a * sv(x, ip[x] + 1, jp[x]) - ip[x] * sv(x, ip[x] - 1, jp[x]);
                const double dsB = 2.0 * b * sv(x, ip[x], jp[x] + 1) - jp[x] * sv(x, ip[x], jp[x] - 1);
                const double dtA = 2.0 * a * tv(x, ip[x] + 1, jp[x]) - ip[x] * tv(x, ip[x] - 1, jp[x]);
                const double dtB = 2.0 * b * tv(x, ip[x], jp[x] + 1) - jp[x] * tv(x, ip[x], jp[x] - 1);
                const double SA = dsA * sy * sz, SB = dsB * sy * sz;
                const double TA = dtA * sy * sz + dsA * (ty * sz + sy * tz);
                const double TB = dtB * sy * sz + dsB * (ty * sz + sy * tz);
                gA0[x] += pre * nrm * (dv * TA - wv * SA);
                gB0[x] += pre * nrm * (dv * TB - wv * SB);
            }
        }
    for (int x = 0; x < 3; ++x) {
        add(q.atA, x, gA0[x]);
        add(q.atB, x, gB0[x]);
    }
}

// Nuclear attraction terms of nucleus cnuc for the primitive pair: one Rys rule and one set of 1-D
// tables (with one extra power on each side), shared by all component pairs.
template <typename Add>
QCG_HD void onee_grad_nuclear(const BasisView& B, int si, int sj, int pi, int pj, int cnuc, const double* D,
                              const double* rys_tab, Add&& add) {
    constexpr int LV = kMaxL + 2;                 // table dimension for i, j <= l + 1
    constexpr int kNr = (2 * kMaxL + 1) / 2 + 1;  // nodes for la = lb = kMaxL
    const OneePair q = onee_pair_setup(B, si, sj, pi, pj);
    const double a = q.a, b = q.b;
    const int nr = (q.la + q.lb + 1) / 2 + 1;
    const int nmax = q.la + q.lb + 2;
    const double* C = &B.xyz[3 * cnuc];
    const double PC[3] = {q.P[0] - C[0], q.P[1] - C[1], q.P[2] - C[2]};
    const double Tn = q.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
    double rx[QCG_RYS_NMAX], rw[QCG_RYS_NMAX];
    onee_rys(nr, Tn, rys_tab, rx, rw);
    double g[kNr][3][LV * LV];
    for (int r = 0; r < nr; ++r) {
        const double u = rx[r];
        for (int x = 0; x < 3; ++x) {
            double h[2 * kMaxL + 3];
            const double c0 = q.PA[x] - u * PC[x];
            const double bb = (1.0 - u) * q.oo2p;
            h[0] = 1.0;
            if (nmax > 0) h[1] = c0;
            for (int n = 1; n < nmax; ++n) h[n + 1] = c0 * h[n] + n * bb * h[n - 1];
            double t2[(2 * kMaxL + 3) * LV];
            for (int n = 0; n <= nmax; ++n) t2[n * LV] = h[n];
            for (int j = 1; j <= q.lb + 1; ++j)
                for (int n = 0; n + j <= nmax; ++n)
                    t2[n * LV + j] = t2[(n + 1) * LV + j - 1] + q.AB[x] * t2[n * LV + j - 1];
            for (int i = 0; i <= q.la + 1; ++i)
                for (int j = 0; j <= q.lb + 1; ++j) g[r][x][i * LV + j] = (i + j <= nmax) ? t2[i * LV + j] : 0.0;
        }
    }
    const double pz = -(double)B.Z[cnuc] * 2.0 * kPi / q.p * q.kab * q.ca * q.cb;
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    for (int ia = 0; ia < q.na; ++ia)
        for (int jb = 0; jb < q.nb; ++jb) {
            const double dv = D[(size_t)(q.bfa + ia) * B.nbf + q.bfb + jb];
            if (dv == 0.0) continue;
            const int ip[3] = {cart_x(q.la, ia), cart_y(q.la, ia), cart_z(q.la, ia)};
            const int jp[3] = {cart_x(q.lb, jb), cart_y(q.lb, jb), cart_z(q.lb, jb)};
            const double pv = pz * onee_norm_comp(q.la, ia) * onee_norm_comp(q.lb, jb) * dv;
            for (int r = 0; r < nr; ++r) {
                auto gv = [&](int x, int i, int j) { return (i < 0 || j < 0) ? 0.0 : g[r][x][i * LV + j]; };
                for (int x = 0; x < 3; ++x) {
                    const int y = (x + 1) % 3, z = (x + 2) % 3;
                    const double oth = rw[r] * pv * gv(y, ip[y], jp[y]) * gv(z, ip[z], jp[z]);
                    gA[x] += oth * (2.0 * a * gv(x, ip[x] + 1, jp[x]) - ip[x] * gv(x, ip[x] - 1, jp[x]));
                    gB[x] += oth * (2.0 * b * gv(x, ip[x], jp[x] + 1) - jp[x] * gv(x, ip[x], jp[x] - 1));
                }
            }
        }
    for (int x = 0; x < 3; ++x) {
        add(q.atA, x, gA[x]);
        add(q.atB, x, gB[x]);
        add(cnuc, x, -(gA[x] + gB[x]));
    }
}

// One term of S, T, V for the Cartesian component pair (ia, jb) of the ordered shell pair (si, sj)
// and its primitive pair (pi, pj) (overlap_matrix, integrals/overlap.rs:647; kinetic,
// integrals/kinetic.rs; nuclear attraction, integrals/nuclear.rs:226), component norms left out:
// cnuc < 0: overlap and kinetic into out[0], out[1]; cnuc >= 0: the attraction to nucleus cnuc
// into out[2].  Overlap and kinetic from the Obara-Saika 1-D tables; V by Rys quadrature, the 1-D
// moments h_n = <(x-A)^n> by their two-term recurrence and (x-B)^j moved onto A binomially.
QCG_HD void onee_element_term(const BasisView& B, int si, int sj, int ia, int jb, int pi, int pj, int cnuc,
                              const double* rys_tab, double* out) {
    constexpr int LD = kMaxL + 3;
    const int la = B.l[si], lb = B.l[sj];
    const int ip[3] = {cart_x(la, ia), cart_y(la, ia), cart_z(la, ia)};
    const int jp[3] = {cart_x(lb, jb), cart_y(lb, jb), cart_z(lb, jb)};
    const OneePair q = onee_pair_setup(B, si, sj, pi, pj);
    if (cnuc < 0) {
        const double a = q.a, b = q.b;
        double s1[LD * LD];
        double sx[3], tx[3];
        for (int x = 0; x < 3; ++x) {
            const int i = ip[x], j = jp[x];
            onee_overlap_1d(q.PA[x], q.PB[x], q.oo2p, i + 1, j + 1, s1, LD);
            sx[x] = s1[i * LD + j];
            double v = 4.0 * a * b * s1[(i + 1) * LD + j + 1];
            if (i > 0) v -= 2.0 * b * i * s1[(i - 1) * LD + j + 1];
            if (j > 0) v -= 2.0 * a * j * s1[(i + 1) * LD + j - 1];
            if (i > 0 && j > 0) v += (double)(i * j) * s1[(i - 1) * LD + j - 1];
            tx[x] = 0.5 * v;
        }
        const double pre_s = q.kab * pow(kPi / q.p, 1.5) * q.ca * q.cb;
        out[0] += pre_s * sx[0] * sx[1] * sx[2];
        out[1] += pre_s * (tx[0] * sx[1] * sx[2] + sx[0] * tx[1] * sx[2] + sx[0] * sx[1] * tx[2]);
        return;
    }
    const int nr = (la + lb) / 2 + 1;
    const double* C = &B.xyz[3 * cnuc];
    const double PC[3] = {q.P[0] - C[0], q.P[1] - C[1], q.P[2] - C[2]};
    const double Tn = q.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
    double rx[QCG_RYS_NMAX], rw[QCG_RYS_NMAX];
    onee_rys(nr, Tn, rys_tab, rx, rw);
    double acc = 0.0;
    for (int r = 0; r < nr; ++r) {
        const double u = rx[r];
        double prod = rw[r];
        for (int x = 0; x < 3; ++x) {
            const int i = ip[x], j = jp[x];
            double h[2 * kMaxL + 1];
            const double c0 = q.PA[x] - u * PC[x];
            const double bb = (1.0 - u) * q.oo2p;
            h[0] = 1.0;
            if (i + j > 0) h[1] = c0;
            for (int n = 1; n < i + j; ++n) h[n + 1] = c0 * h[n] + n * bb * h[n - 1];
            // (x-B)^j = sum_k C(j,k) (x-A)^k (A-B)^(j-k), from k = j down
            double g = 0.0, pw = 1.0, cb = 1.0;
            for (int k = j; k >= 0; --k) {
                g += cb * pw * h[i + k];
                pw *= q.AB[x];
                cb = cb * k / (j - k + 1);
            }
            prod *= g;
        }
        acc += prod;
    }
    out[2] += -(double)B.Z[cnuc] * 2.0 * kPi / q.p * q.kab * q.ca * q.cb * acc;
}

}  // namespace qcg
