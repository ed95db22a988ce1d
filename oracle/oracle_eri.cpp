// CPU ORACLE — test infrastructure only.
//
// A plain FP64 restatement of qc-core's two-electron hot path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  Nothing
// in the product links or calls this file.
//
// Each function cites the reference function it restates (paths relative to the
// ExaPsi/IQCP checkout).  Parity status of each function (what pins it) is
// recorded next to it.
//
//   or_boys_all        boys/mod.rs:587 (+:481 series, :557 erf/upward)      pinned: tests/golden/boys (scipy hyp1f1, 1e-12) + mpmath
//   or_rys_roots       rys/mod.rs:253 -> :468 rdk, :585 schmidt, :687 roots,  pinned: moment reconstruction (Gauss uniqueness) + tests/golden/rys
//                      :801 hessenberg_qr, :881 qr_step, :953 bisection
//   or_shell_eri       integrals/eri/mod.rs:458-854                           pinned: PySCF STO-3G tensors (content/presets) + (ss|ss) closed form
//   or_schwarz         integrals/eri/mod.rs:1263-1290                         pinned: via or_shell_eri
//   or_eri_compressed  integrals/eri/mod.rs:992-1070                          pinned: PySCF STO-3G tensors
//   or_build_fock      scf/mod.rs:825-881                                     pinned: H2 PySCF J/K/F (scf/fock.rs:192-229) + numpy einsum
//   or_build_jk        dft/ks_scf.rs:1306 / :1342 / :1369                     pinned: numpy einsum on the full tensor
//   or_jk_direct       eri_compressed + build_jk without the stored tensor     pinned: equals or_build_jk on or_eri_compressed
//   or_one_electron    integrals/overlap.rs:647, kinetic.rs, nuclear.rs:226   pinned: PySCF S and h_core of the STO-3G presets
//                      (binomial Gaussian moments; V by Rys over or_rys_roots)  (content/presets), closed-form s-s values
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

extern "C" {

// ---------------------------------------------------------------- Boys
static const double kTurnover[40] = {
    0.0, 0.0, 0.866025403784, 1.295010032056, 1.705493613097, 2.106432965305,
    2.501471934009, 2.892473348218, 3.280525047072, 3.666320693281, 4.05033123037,
    4.432891808508, 4.814249856864, 5.194593501454, 5.574069276051, 5.952793645111,
    6.330860773135, 6.708347923415, 7.08531930745, 7.461828891625, 7.837922483937,
    8.213639312398, 8.589013237349, 8.964073695432, 9.338846443746, 9.713354153046,
    10.08761688545, 10.46165248270, 10.83547688448, 11.20910439128, 11.58254788331,
    11.95581900374, 12.32892831326, 12.70188542111, 13.07469909673, 13.44737736550,
    13.81992759110, 14.19235654675, 14.56467047710, 14.93687515212};

static double turnover(int m) {
    if (m < 40) return kTurnover[m];
    return 14.93687515212 + 0.372 * (double)(m - 39);
}

// boys_all_internal (boys/mod.rs:587-621): regime chosen by the turnover of m_max.
int or_boys_all(int m, double t, double* f) {
    if (m < 0 || m > 50) return 1;
    if (t < 0.0) return 1;
    if (t < 1e-14) {
        for (int i = 0; i <= m; ++i) f[i] = 1.0 / (2.0 * i + 1.0);
        return 0;
    }
    if (t < turnover(m)) {
        // boys_series_internal (:481-527)
        double b = m + 0.5;
        double e = 0.5 * std::exp(-t);
        double x = e, s = e;
        double tol = (2.220446049250313e-16 * 0.5) * e;
        double bi = b + 1.0;
        while (x > tol) {
            x *= t / bi;
            s += x;
            bi += 1.0;
        }
        f[m] = s / b;
        double bc = b;
        for (int i = m; i >= 1; --i) {
            bc -= 1.0;
            f[i - 1] = (e + t * f[i]) / bc;
        }
    } else {
        // boys_recurrence_internal (:557-578); erf = fdlibm erf (libm crate, :649-651)
        double tt = std::sqrt(t);
        f[0] = 0.8862269254527580136490837416705725913987747280611935641069038949264 / tt * std::erf(tt);
        double e = std::exp(-t);
        double b = 0.5 / t;
        for (int i = 1; i <= m; ++i) f[i] = b * ((2 * i - 1) * f[i - 1] - e);
    }
    return 0;
}

// ---------------------------------------------------------------- Rys (RDK)
enum { MAXR = 10 };
static const double kThreshZero = 2.220446049250313e-16 * 8.0;  // THRESHOLD_ZERO (rys/mod.rs:50)

static int schmidt(double* cs, const double* mom, int n) {
    double v[MAXR + 1];
    cs[0] = 1.0 / std::sqrt(mom[0]);
    double fac = -mom[1] / mom[0];
    double tmp = mom[2] + fac * mom[1];
    if (tmp <= 0.0) return 1;
    double tis = 1.0 / std::sqrt(tmp);
    cs[n] = fac * tis;
    cs[1 + n] = tis;
    for (int j = 2; j < n; ++j) {
        for (int k = 0; k < j; ++k) v[k] = 0.0;
        double fc = mom[j + j];
        for (int k = 0; k < j; ++k) {
            double dot = 0.0;
            for (int i = 0; i <= k; ++i) dot += cs[i + k * n] * mom[i + j];
            for (int i = 0; i <= k; ++i) v[i] -= dot * cs[i + k * n];
            fc -= dot * dot;
        }
        if (fc <= 0.0) {
            if (fc == 0.0) {
                for (int jj = j; jj < n; ++jj)
                    for (int ii = 0; ii <= jj; ++ii) cs[ii + jj * n] = 0.0;
                return 0;
            }
            return 1;
        }
        double fis = 1.0 / std::sqrt(fc);
        cs[j + j * n] = fis;
        for (int k = 0; k < j; ++k) cs[k + j * n] = fis * v[k];
    }
    return 0;
}

static void qr_step(double* a, int nr, int n0, int n1, double shift) {
    int m1 = n0 + 1;
    double cv = a[n0 * nr + n0] - shift;
    double sv = a[m1 * nr + n0];
    double v = std::sqrt(cv * cv + sv * sv);
    double c = 1.0, s = 0.0;
    if (v != 0.0) { c = cv / v; s = sv / v; }
    for (int k = n0; k < nr; ++k) {
        double x = a[n0 * nr + k], y = a[m1 * nr + k];
        a[n0 * nr + k] = c * x + s * y;
        a[m1 * nr + k] = c * y - s * x;
    }
    int m3 = std::min(n1, n0 + 3);
    for (int k = 0; k < m3; ++k) {
        double x = a[k * nr + n0], y = a[k * nr + m1];
        a[k * nr + n0] = c * x + s * y;
        a[k * nr + m1] = c * y - s * x;
    }
    for (int j = n0; j < n1 - 2; ++j) {
        int j1 = j + 1, j2 = j + 2;
        double cv2 = a[j1 * nr + j], sv2 = a[j2 * nr + j];
        v = std::sqrt(cv2 * cv2 + sv2 * sv2);
        a[j1 * nr + j] = v;
        a[j2 * nr + j] = 0.0;
        c = 1.0; s = 0.0;
        if (v != 0.0) { c = cv2 / v; s = sv2 / v; }
        for (int k = j1; k < nr; ++k) {
            double x = a[j1 * nr + k], y = a[j2 * nr + k];
            a[j1 * nr + k] = c * x + s * y;
            a[j2 * nr + k] = c * y - s * x;
        }
        int m4 = std::min(n1, j + 4);
        for (int k = 0; k < m4; ++k) {
            double x = a[k * nr + j1], y = a[k * nr + j2];
            a[k * nr + j1] = c * x + s * y;
            a[k * nr + j2] = c * y - s * x;
        }
    }
}

static int hessenberg_qr(double* a, int nr) {
    const double eps = 1e-15;  // QR_EPSILON (rys/mod.rs:62)
    const int maxit = 30;      // QR_MAX_ITERATIONS (rys/mod.rs:58)
    int n0 = 0, n1 = nr, its = 0;
    for (int ic = 0; ic < nr * maxit; ++ic) {
        int k = n0;
        while (k + 1 < n1) {
            double s = std::fabs(a[k * nr + k]) + std::fabs(a[(k + 1) * nr + k + 1]);
            if (std::fabs(a[(k + 1) * nr + k]) < eps * s) break;
            ++k;
        }
        int k1 = k + 1;
        if (k1 < n1) {
            a[k1 * nr + k] = 0.0;
            n0 = k1;
            its = 0;
            if (n0 + 1 >= n1) {
                n0 = 0;
                n1 = k1;
                if (n1 < 2) return 0;
            }
        } else {
            int m1 = n1 - 1, m2 = n1 - 2;
            double a11 = a[m1 * nr + m1], a22 = a[m2 * nr + m2];
            double t = a11 + a22;
            double s = (a11 - a22) * (a11 - a22);
            s += 4.0 * a[m1 * nr + m2] * a[m2 * nr + m1];
            double shift;
            if (s > 0.0) {
                double ss = std::sqrt(s);
                double aa = (t + ss) * 0.5, bb = (t - ss) * 0.5;
                shift = (std::fabs(a11 - aa) > std::fabs(a11 - bb)) ? bb : aa;
            } else {
                if (n1 == 2) return 1;
                shift = t * 0.5;
            }
            ++its;
            qr_step(a, nr, n0, n1, shift);
            if (its > maxit) return 1;
        }
    }
    return 1;
}

static double poly_eval(const double* a, int order, double x) {
    double p = a[order];
    for (int i = 1; i <= order; ++i) p = p * x + a[order - i];
    return p;
}

static int bisection_refine(const double* cs, double* roots, int order, int n1) {
    const double* a = cs + order * n1;
    double x1init = 0.0, p1init = a[0];
    for (int m = 0; m < order; ++m) {
        double x0i = x1init, p0i = p1init;
        x1init = roots[m];
        p1init = poly_eval(a, order, x1init);
        if (p1init == 0.0) continue;
        if (p0i * p1init > 0.0) return 1;
        double x0, x1, p0, p1;
        if (x0i <= x1init) { x0 = x0i; x1 = x1init; p0 = p0i; p1 = p1init; }
        else { x0 = x1init; x1 = x0i; p0 = p1init; p1 = p0i; }
        if (p1 == 0.0) { roots[m] = x1; continue; }
        if (p0 == 0.0) { roots[m] = x0; continue; }
        double xi = x0 + (x0 - x1) / (p1 - p0) * p0;
        for (int it = 0; it < 200; ++it) {  // BISECTION_MAX_ITERATIONS (rys/mod.rs:66)
            if (std::fabs(x1 - x0) <= x1 * 1e-15) break;  // ROOT_ACCURACY (rys/mod.rs:54)
            double pi = poly_eval(a, order, xi);
            if (pi == 0.0) break;
            if (p0 * pi <= 0.0) { x1 = xi; p1 = pi; xi = x0 * 0.25 + xi * 0.75; }
            else { x0 = xi; p0 = pi; xi = xi * 0.75 + x1 * 0.25; }
            pi = poly_eval(a, order, xi);
            if (pi == 0.0) break;
            if (p0 * pi <= 0.0) { x1 = xi; p1 = pi; }
            else { x0 = xi; p0 = pi; }
            xi = x0 + (x0 - x1) / (p1 - p0) * p0;
        }
        roots[m] = xi;
    }
    return 0;
}

static int poly_roots(double* roots, const double* cs, int nr, int n1) {
    if (nr == 1) { roots[0] = -cs[n1] / cs[1 + n1]; return 0; }
    if (nr == 2) {
        double a = cs[2 + 2 * n1], b = cs[1 + 2 * n1], c = cs[2 * n1];
        double disc = b * b - 4.0 * a * c;
        if (disc < 0.0) return 1;
        double sd = std::sqrt(disc);
        roots[0] = (-b - sd) / (2.0 * a);
        roots[1] = (-b + sd) / (2.0 * a);
        return 0;
    }
    double A[MAXR * MAXR];
    std::memset(A, 0, sizeof(A));
    int off = nr * n1;
    double lead = cs[nr + off];
    if (std::fabs(lead) < kThreshZero) return 1;
    double fac = -1.0 / lead;
    for (int i = 0; i < nr; ++i) A[nr - 1 - i] = cs[off + i] * fac;
    for (int i = 0; i < nr - 1; ++i) A[(i + 1) * nr + i] = 1.0;
    if (hessenberg_qr(A, nr) == 0) {
        for (int i = 0; i < nr; ++i) roots[nr - 1 - i] = A[i * nr + i];
    } else {
        double a = cs[2 + 2 * n1], b = cs[1 + 2 * n1], c = cs[2 * n1];
        double disc = b * b - 4.0 * a * c;
        double sd = disc >= 0.0 ? std::sqrt(disc) : 0.0;
        roots[0] = 0.5 * (-b - sd) / a;
        roots[1] = 0.5 * (-b + sd) / a;
        for (int k = 2; k < nr; ++k) roots[k] = 1.0;
        for (int k = 2; k < nr; ++k)
            if (bisection_refine(cs, roots, k + 1, n1)) return 1;
    }
    return 0;
}

// rys_roots (rys/mod.rs:253) + rdk_rys_roots (:468).  Roots are u = t^2 in [0,1).
// Returns 0 on success, nonzero where the reference returns Err.
int or_rys_roots(int nroots, double t, double* roots, double* weights) {
    if (nroots < 1 || nroots > MAXR || t < 0.0) return 1;
    double mom[2 * MAXR + 2];
    if (or_boys_all(2 * nroots, t, mom)) return 1;
    if (std::fabs(mom[0]) < kThreshZero) {
        for (int i = 0; i < nroots; ++i) roots[i] = weights[i] = 0.0;
        return 0;
    }
    if (nroots == 1) {
        double ratio = mom[1] / (mom[0] - mom[1]);
        roots[0] = ratio >= 0.0 ? ratio / (1.0 + ratio) : 0.0;
        weights[0] = mom[0];
        return 0;
    }
    int n1 = nroots + 1;
    double cs[(MAXR + 1) * (MAXR + 1)];
    std::memset(cs, 0, sizeof(cs));
    if (schmidt(cs, mom, n1)) return 2;
    double r[MAXR], w[MAXR];
    if (poly_roots(r, cs, nroots, n1)) return 3;
    for (int k = 0; k < nroots; ++k) {
        double root = r[k];
        if (std::fabs(root - 1.0) < kThreshZero) { r[k] = 0.0; w[k] = 0.0; continue; }
        double dum = 1.0 / mom[0];
        for (int j = 1; j < nroots; ++j) {
            double p = poly_eval(cs + j * n1, j, root);
            dum += p * p;
        }
        w[k] = 1.0 / dum;
    }
    // RysResult::validate (rys/mod.rs:190-204): roots in [0,1), weights >= 0
    for (int k = 0; k < nroots; ++k) {
        if (!(r[k] >= 0.0 && r[k] < 1.0)) return 4;
        if (w[k] < 0.0) return 4;
    }
    for (int k = 0; k < nroots; ++k) { roots[k] = r[k]; weights[k] = w[k]; }
    return 0;
}

// ---------------------------------------------------------------- basis view
struct OrBasis {
    int nshell;
    const int* l;
    const int* nprim;
    const int* prim_off;
    const double* exps;
    const double* coefs;
    const double* centers;  // 3*nshell
    const int* bf_off;
    int nbf;
};

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static void cart_comps(int l, int (*c)[3]) {
    int n = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) { c[n][0] = i; c[n][1] = j; c[n][2] = l - i - j; ++n; }
}

static double dfact(int n) {  // (n)!! for odd n, 1 for n <= 1
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

// cartesian_gaussian_normalization (integrals/overlap.rs:343-377)
static double cnorm(double alpha, const int* p) {
    const double PI = 3.14159265358979323846;
    double two_a_pi = 2.0 * alpha / PI;
    double base = std::sqrt(std::sqrt(two_a_pi * two_a_pi * two_a_pi));
    int L = p[0] + p[1] + p[2];
    if (L == 0) return base;
    double fa = 4.0 * alpha, num;
    switch (L) {
        case 1: num = std::sqrt(fa); break;
        case 2: num = fa; break;
        case 3: num = fa * std::sqrt(fa); break;
        case 4: num = fa * fa; break;
        default: num = std::pow(fa, L / 2.0);
    }
    double den = std::sqrt(dfact(p[0] > 0 ? 2 * p[0] - 1 : 0) * dfact(p[1] > 0 ? 2 * p[1] - 1 : 0) *
                           dfact(p[2] > 0 ? 2 * p[2] - 1 : 0));
    return base * num / den;
}

// build_2d_into (integrals/eri/vrr_2d.rs:143-213)
static void vrr2d(double* g, int nb, int nk, double c00, double c0p, double b00, double b10, double b01) {
    int cols = nk + 1, size = (nb + 1) * cols;
    for (int i = 0; i < size; ++i) g[i] = 0.0;
    g[0] = 1.0;
    if (nb == 0 && nk == 0) return;
    if (nb > 0) {
        g[cols] = c00;
        for (int n = 1; n < nb; ++n) g[(n + 1) * cols] = c00 * g[n * cols] + n * b10 * g[(n - 1) * cols];
    }
    for (int m = 0; m < nk; ++m) {
        g[m + 1] = (m == 0) ? c0p * g[0] : c0p * g[m] + m * b01 * g[m - 1];
        for (int n = 1; n <= nb; ++n) {
            double gnm = g[n * cols + m];
            double gnm1 = m > 0 ? g[n * cols + m - 1] : 0.0;
            double gn1m = g[(n - 1) * cols + m];
            g[n * cols + m + 1] = c0p * gnm + m * b01 * gnm1 + n * b00 * gn1m;
        }
    }
}

// horizontal_transfer_1d (integrals/eri/htr_4d.rs:70-103, recursive form :257)
static double htr(const double* g, int nk, int i, int j, int k, int l, double ab, double cd) {
    if (j == 0 && l == 0) return g[i * (nk + 1) + k];
    if (l > 0) return htr(g, nk, i, j, k + 1, l - 1, ab, cd) + cd * htr(g, nk, i, j, k, l - 1, ab, cd);
    return htr(g, nk, i + 1, j - 1, k, l, ab, cd) + ab * htr(g, nk, i, j - 1, k, l, ab, cd);
}

// shell_eri (integrals/eri/mod.rs:458-854).  out: row-major [i][j][k][l].
static void shell_eri_impl(const OrBasis& B, int si, int sj, int sk, int sl, double* out) {
    const double PI = 3.14159265358979323846;
    int sh[4] = {si, sj, sk, sl};
    int L[4], N[4];
    int comps[4][6][3];
    for (int a = 0; a < 4; ++a) {
        L[a] = B.l[sh[a]];
        N[a] = ncart(L[a]);
        cart_comps(L[a], comps[a]);
    }
    int nq = N[0] * N[1] * N[2] * N[3];
    for (int i = 0; i < nq; ++i) out[i] = 0.0;
    const double* A = B.centers + 3 * si;
    const double* Bc = B.centers + 3 * sj;
    const double* C = B.centers + 3 * sk;
    const double* D = B.centers + 3 * sl;
    double ab[3] = {A[0] - Bc[0], A[1] - Bc[1], A[2] - Bc[2]};
    double cd[3] = {C[0] - D[0], C[1] - D[1], C[2] - D[2]};
    double ab2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    double cd2 = cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2];
    int ltot = L[0] + L[1] + L[2] + L[3];
    int nroots = ltot / 2 + 1;
    int nb = L[0] + L[1], nk = L[2] + L[3];
    int vsz = (nb + 1) * (nk + 1);
    std::vector<double> gx(vsz), gy(vsz), gz(vsz);
    int np[4];
    const double *ex[4], *co[4];
    for (int a = 0; a < 4; ++a) {
        np[a] = B.nprim[sh[a]];
        ex[a] = B.exps + B.prim_off[sh[a]];
        co[a] = B.coefs + B.prim_off[sh[a]];
    }
    for (int pi = 0; pi < np[0]; ++pi)
        for (int pj = 0; pj < np[1]; ++pj) {
            double ai = ex[0][pi], aj = ex[1][pj];
            double kij = std::exp(-(ai * aj / (ai + aj)) * ab2);
            if (kij < 1e-15) continue;
            for (int pk = 0; pk < np[2]; ++pk)
                for (int pl = 0; pl < np[3]; ++pl) {
                    double ak = ex[2][pk], al = ex[3][pl];
                    double kkl = std::exp(-(ak * al / (ak + al)) * cd2);
                    if (kij * kkl < 1e-15) continue;
                    // GaussianProduct2e::new_prescreened (gaussian_product_2e.rs:260-342)
                    double p = ai + aj, q = ak + al;
                    double P[3], Q[3], pa[3], qc[3], pq[3];
                    for (int x = 0; x < 3; ++x) {
                        P[x] = (ai * A[x] + aj * Bc[x]) / p;
                        Q[x] = (ak * C[x] + al * D[x]) / q;
                        pa[x] = P[x] - A[x];
                        qc[x] = Q[x] - C[x];
                        pq[x] = P[x] - Q[x];
                    }
                    double pq2 = pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2];
                    double rho = p * q / (p + q);
                    double T = rho * pq2;
                    double pref = 2.0 * PI * PI * std::sqrt(PI) / (p * q * std::sqrt(p + q)) * kij * kkl;
                    double cbase = co[0][pi] * co[1][pj] * co[2][pk] * co[3][pl];
                    double rr[MAXR], ww[MAXR];
                    int nr_eff = nroots;
                    bool tzero = false;
                    if (or_rys_roots(nroots, T, rr, ww)) {
                        if (T < 1e-15) { tzero = true; nr_eff = 1; }
                        else continue;  // reference drops the primitive quartet (mod.rs:736)
                    }
                    for (int r = 0; r < nr_eff; ++r) {
                        double b00, b10, b01, c00[3], c0p[3], wpref;
                        if (tzero) {
                            // RysCoefficients::compute_t_zero (rys_coefficients.rs:183-217)
                            double tmp4 = 0.5 / (p * q);
                            b00 = 0.0; b10 = tmp4 * q; b01 = tmp4 * p;
                            for (int x = 0; x < 3; ++x) { c00[x] = pa[x]; c0p[x] = qc[x]; }
                            wpref = pref * cbase;
                        } else {
                            // RysCoefficients::compute (rys_coefficients.rs:94-180)
                            double u = rr[r];
                            double ut = (u < 1.0 - 1e-15) ? u / (1.0 - u) : 1e15;
                            double u2 = rho * ut;
                            double tmp4 = 0.5 / (u2 * (p + q) + p * q);
                            double tmp5 = u2 * tmp4;
                            b00 = tmp5; b10 = tmp5 + tmp4 * q; b01 = tmp5 + tmp4 * p;
                            double t2 = 2.0 * tmp5 * q, t3 = 2.0 * tmp5 * p;
                            for (int x = 0; x < 3; ++x) { c00[x] = pa[x] - t2 * pq[x]; c0p[x] = qc[x] + t3 * pq[x]; }
                            wpref = pref * ww[r] * cbase;
                        }
                        vrr2d(gx.data(), nb, nk, c00[0], c0p[0], b00, b10, b01);
                        vrr2d(gy.data(), nb, nk, c00[1], c0p[1], b00, b10, b01);
                        vrr2d(gz.data(), nb, nk, c00[2], c0p[2], b00, b10, b01);
                        for (int ii = 0; ii < N[0]; ++ii) {
                            const int* ci = comps[0][ii];
                            double ni = cnorm(ai, ci);
                            for (int jj = 0; jj < N[1]; ++jj) {
                                const int* cj = comps[1][jj];
                                double nj = cnorm(aj, cj);
                                for (int kk = 0; kk < N[2]; ++kk) {
                                    const int* ck = comps[2][kk];
                                    double nkk = cnorm(ak, ck);
                                    for (int ll = 0; ll < N[3]; ++ll) {
                                        const int* cl = comps[3][ll];
                                        double nl = cnorm(al, cl);
                                        double ix = htr(gx.data(), nk, ci[0], cj[0], ck[0], cl[0], ab[0], cd[0]);
                                        double iy = htr(gy.data(), nk, ci[1], cj[1], ck[1], cl[1], ab[1], cd[1]);
                                        double iz = htr(gz.data(), nk, ci[2], cj[2], ck[2], cl[2], ab[2], cd[2]);
                                        int idx = ((ii * N[1] + jj) * N[2] + kk) * N[3] + ll;
                                        out[idx] += wpref * ni * nj * nkk * nl * ix * iy * iz;
                                    }
                                }
                            }
                        }
                    }
                }
        }
}

static OrBasis make_basis(int nshell, const int* l, const int* nprim, const int* prim_off,
                          const double* exps, const double* coefs, const double* centers,
                          const int* bf_off, int nbf) {
    OrBasis b{nshell, l, nprim, prim_off, exps, coefs, centers, bf_off, nbf};
    return b;
}

int or_shell_eri(int nshell, const int* l, const int* nprim, const int* prim_off, const double* exps,
                 const double* coefs, const double* centers, const int* bf_off, int nbf,
                 int si, int sj, int sk, int sl, double* out) {
    OrBasis B = make_basis(nshell, l, nprim, prim_off, exps, coefs, centers, bf_off, nbf);
    for (int s : {si, sj, sk, sl}) {
        if (s < 0 || s >= nshell) return 1;
        if (l[s] > 2) return 2;
    }
    shell_eri_impl(B, si, sj, sk, sl, out);
    return 0;
}

// compute_schwarz_bounds (integrals/eri/mod.rs:1263-1290): Q[i][j] = sqrt(max |(ij|ij)|)
int or_schwarz(int nshell, const int* l, const int* nprim, const int* prim_off, const double* exps,
               const double* coefs, const double* centers, const int* bf_off, int nbf, double* Q) {
    OrBasis B = make_basis(nshell, l, nprim, prim_off, exps, coefs, centers, bf_off, nbf);
    std::vector<double> blk(1296);
    for (int i = 0; i < nshell; ++i)
        for (int j = 0; j <= i; ++j) {
            shell_eri_impl(B, i, j, i, j, blk.data());
            int n = ncart(l[i]) * ncart(l[j]);
            n = n * n;
            double mx = 0.0;
            for (int k = 0; k < n; ++k) mx = std::max(mx, std::fabs(blk[k]));
            double b = std::sqrt(mx);
            Q[i * nshell + j] = b;
            Q[j * nshell + i] = b;
        }
    return 0;
}

static inline int64_t pidx(int64_t i, int64_t j) {
    if (i < j) std::swap(i, j);
    return i * (i + 1) / 2 + j;
}
static inline int64_t eidx(int64_t i, int64_t j, int64_t k, int64_t l) {
    int64_t a = pidx(i, j), b = pidx(k, l);
    if (a < b) std::swap(a, b);
    return a * (a + 1) / 2 + b;
}

// eri_compressed (integrals/eri/mod.rs:992-1070).  Also reports the number of shell quartets
// that passed the Schwarz test (loop order si>=sj, all sk, sl<=sk, exactly as the reference).
int or_eri_compressed(int nshell, const int* l, const int* nprim, const int* prim_off, const double* exps,
                      const double* coefs, const double* centers, const int* bf_off, int nbf,
                      double* eri, int64_t* n_computed) {
    OrBasis B = make_basis(nshell, l, nprim, prim_off, exps, coefs, centers, bf_off, nbf);
    int64_t npair = (int64_t)nbf * (nbf + 1) / 2;
    int64_t neri = npair * (npair + 1) / 2;
    for (int64_t i = 0; i < neri; ++i) eri[i] = 0.0;
    std::vector<double> Q((size_t)nshell * nshell);
    or_schwarz(nshell, l, nprim, prim_off, exps, coefs, centers, bf_off, nbf, Q.data());
    std::vector<double> blk(1296);
    int64_t cnt = 0;
    for (int si = 0; si < nshell; ++si)
        for (int sj = 0; sj <= si; ++sj) {
            double qij = Q[si * nshell + sj];
            for (int sk = 0; sk < nshell; ++sk)
                for (int sl = 0; sl <= sk; ++sl) {
                    double qkl = Q[sk * nshell + sl];
                    if (qij * qkl < 1e-12) continue;
                    ++cnt;
                    shell_eri_impl(B, si, sj, sk, sl, blk.data());
                    int ni = ncart(l[si]), nj = ncart(l[sj]), nk = ncart(l[sk]), nl = ncart(l[sl]);
                    for (int a = 0; a < ni; ++a)
                        for (int b = 0; b < nj; ++b) {
                            int i = bf_off[si] + a, j = bf_off[sj] + b;
                            if (i < j) continue;
                            for (int c = 0; c < nk; ++c)
                                for (int d = 0; d < nl; ++d) {
                                    int k = bf_off[sk] + c, ll = bf_off[sl] + d;
                                    if (k < ll) continue;
                                    eri[eidx(i, j, k, ll)] = blk[((a * nj + b) * nk + c) * nl + d];
                                }
                        }
                }
        }
    if (n_computed) *n_computed = cnt;
    return 0;
}

// build_fock (scf/mod.rs:825-881).  H, D, F are nbf x nbf (symmetric; layout-agnostic).
int or_build_fock(const double* H, const double* D, const double* eri, int nbf, double* F) {
    for (int i = 0; i < nbf * nbf; ++i) F[i] = H[i];
    for (int mu = 0; mu < nbf; ++mu)
        for (int nu = 0; nu <= mu; ++nu) {
            double g = 0.0;
            for (int la = 0; la < nbf; ++la) {
                double dll = D[la * nbf + la];
                g += dll * (eri[eidx(mu, nu, la, la)] - 0.5 * eri[eidx(mu, la, nu, la)]);
                for (int si = 0; si < la; ++si) {
                    double dls = D[la * nbf + si];
                    double j = eri[eidx(mu, nu, la, si)];
                    double k1 = eri[eidx(mu, la, nu, si)];
                    double k2 = eri[eidx(mu, si, nu, la)];
                    g += dls * (2.0 * j - 0.5 * (k1 + k2));
                }
            }
            F[mu * nbf + nu] += g;
            if (mu != nu) F[nu * nbf + mu] += g;
        }
    return 0;
}

// build_jk (dft/ks_scf.rs:1369-1394): full lambda,sigma loop, two lookups per term.
int or_build_jk(const double* D, const double* eri, int nbf, double* J, double* K) {
    for (int mu = 0; mu < nbf; ++mu)
        for (int nu = 0; nu <= mu; ++nu) {
            double j = 0.0, k = 0.0;
            for (int la = 0; la < nbf; ++la)
                for (int si = 0; si < nbf; ++si) {
                    double d = D[la * nbf + si];
                    j += d * eri[eidx(mu, nu, la, si)];
                    k += d * eri[eidx(mu, la, nu, si)];
                }
            J[mu * nbf + nu] = J[nu * nbf + mu] = j;
            K[mu * nbf + nu] = K[nu * nbf + mu] = k;
        }
    return 0;
}

}  // extern "C"

extern "C" {
// J and K straight from the unique shell quartets, without the stored tensor: the integrals of
// or_eri_compressed (same shell_eri, same Schwarz predicate Q_ij Q_kl >= thr,
// integrals/eri/mod.rs:1008-1023) contracted as build_jk does (dft/ks_scf.rs:1369-1394),
// J_pq = sum_rs (pq|rs) D_rs and K_pr = sum_qs (pq|rs) D_qs over the full index range.  Each
// unique basis-function quartet (i >= j, k >= l, pair(ij) >= pair(kl)) is computed once and
// expanded over its distinct index permutations, so every element of the full tensor enters
// exactly once.  Shell quartets run on nthreads OpenMP threads with private J/K, summed at the
// end.  Used where the compressed tensor would not fit a test (decane 6-31G*: 1.43 GB);
// pinned against or_build_jk on the stored tensor.  Returns the number of unique shell quartets
// kept.
int64_t or_jk_direct(int nshell, const int* l, const int* nprim, const int* prim_off, const double* exps,
                     const double* coefs, const double* centers, const int* bf_off, int nbf, const double* D,
                     double thr, int nthreads, double* J, double* K) {
    OrBasis B = make_basis(nshell, l, nprim, prim_off, exps, coefs, centers, bf_off, nbf);
    std::vector<double> Q((size_t)nshell * nshell);
    or_schwarz(nshell, l, nprim, prim_off, exps, coefs, centers, bf_off, nbf, Q.data());
    // unique shell pairs (si >= sj) in pair-index order
    std::vector<int> ps, qs;
    for (int si = 0; si < nshell; ++si)
        for (int sj = 0; sj <= si; ++sj) {
            ps.push_back(si);
            qs.push_back(sj);
        }
    const int np = (int)ps.size();
    const size_t n2 = (size_t)nbf * nbf;
    for (size_t i = 0; i < n2; ++i) J[i] = K[i] = 0.0;
    int64_t kept = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : kept)
    {
        std::vector<double> Jt(n2, 0.0), Kt(n2, 0.0), blk(1296);
#pragma omp for schedule(dynamic, 1)
        for (int P = np - 1; P >= 0; --P) {
            const int si = ps[P], sj = qs[P];
            for (int Qp = 0; Qp <= P; ++Qp) {
                const int sk = ps[Qp], sl = qs[Qp];
                if (Q[si * nshell + sj] * Q[sk * nshell + sl] < thr) continue;
                ++kept;
                shell_eri_impl(B, si, sj, sk, sl, blk.data());
                const int ni = ncart(l[si]), nj = ncart(l[sj]), nk = ncart(l[sk]), nl = ncart(l[sl]);
                for (int a = 0; a < ni; ++a)
                    for (int b = 0; b < nj; ++b) {
                        const int i = bf_off[si] + a, j = bf_off[sj] + b;
                        if (si == sj && i < j) continue;
                        for (int c = 0; c < nk; ++c)
                            for (int d = 0; d < nl; ++d) {
                                const int k = bf_off[sk] + c, ll = bf_off[sl] + d;
                                if (sk == sl && k < ll) continue;
                                if (P == Qp && pidx(i, j) < pidx(k, ll)) continue;
                                const double v = blk[((a * nj + b) * nk + c) * nl + d];
                                int perm[8][4] = {{i, j, k, ll}, {j, i, k, ll}, {i, j, ll, k}, {j, i, ll, k},
                                                  {k, ll, i, j}, {ll, k, i, j}, {k, ll, j, i}, {ll, k, j, i}};
                                for (int t = 0; t < 8; ++t) {
                                    bool dup = false;
                                    for (int u = 0; u < t && !dup; ++u)
                                        dup = perm[u][0] == perm[t][0] && perm[u][1] == perm[t][1] &&
                                              perm[u][2] == perm[t][2] && perm[u][3] == perm[t][3];
                                    if (dup) continue;
                                    const int p = perm[t][0], q = perm[t][1], r = perm[t][2], s = perm[t][3];
                                    Jt[(size_t)p * nbf + q] += D[(size_t)r * nbf + s] * v;
                                    Kt[(size_t)p * nbf + r] += D[(size_t)q * nbf + s] * v;
                                }
                            }
                    }
            }
        }
#pragma omp critical
        for (size_t i = 0; i < n2; ++i) {
            J[i] += Jt[i];
            K[i] += Kt[i];
        }
    }
    return kept;
}

// build_fock (scf/mod.rs:825-881) with the mu loop spread over nthreads OpenMP threads (the
// (mu, nu) elements are independent, so the result is identical to or_build_fock).  Used as
// the timed CPU baseline.
int or_build_fock_omp(const double* H, const double* D, const double* eri, int nbf, double* F, int nthreads) {
    for (int i = 0; i < nbf * nbf; ++i) F[i] = H[i];
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int mu = nbf - 1; mu >= 0; --mu)
        for (int nu = 0; nu <= mu; ++nu) {
            double g = 0.0;
            for (int la = 0; la < nbf; ++la) {
                double dll = D[la * nbf + la];
                g += dll * (eri[eidx(mu, nu, la, la)] - 0.5 * eri[eidx(mu, la, nu, la)]);
                for (int si = 0; si < la; ++si) {
                    double dls = D[la * nbf + si];
                    double j = eri[eidx(mu, nu, la, si)];
                    double k1 = eri[eidx(mu, la, nu, si)];
                    double k2 = eri[eidx(mu, si, nu, la)];
                    g += dls * (2.0 * j - 0.5 * (k1 + k2));
                }
            }
            F[mu * nbf + nu] += g;
            if (mu != nu) F[nu * nbf + mu] += g;
        }
    return 0;
}
}  // extern "C"

extern "C" {
// ---------------------------------------------------------------- one-electron S, T, V
// Written with Gaussian moments instead of recurrences.  Each 1-D factor is the expectation of
// (x-A)^i (x-B)^j under a normal law of mean m and variance v, by binomial expansion about m:
//   overlap: m = P, v = 1/(2p);  nuclear attraction (after the Rys change of variable with root
//   u = t^2): m = P - u (P - C), v = (1 - u)/(2p).
// Kinetic: -1/2 d^2/dx^2 applied to the ket, (x-B)^j e^{-b(x-B)^2} ->
//   j(j-1)(x-B)^{j-2} - 2b(2j+1)(x-B)^j + 4b^2 (x-B)^{j+2}.
static double binom(int n, int k) {
    double r = 1.0;
    for (int t = 1; t <= k; ++t) r = r * (n - k + t) / t;
    return r;
}

static double gauss_moment_1d(int i, int j, double mA, double mB, double v) {
    if (i < 0 || j < 0) return 0.0;
    double s = 0.0;
    for (int k1 = 0; k1 <= i; ++k1)
        for (int k2 = 0; k2 <= j; ++k2) {
            const int k = k1 + k2;
            if (k % 2) continue;
            s += binom(i, k1) * binom(j, k2) * std::pow(mA, i - k1) * std::pow(mB, j - k2) * dfact(k - 1) *
                 std::pow(v, k / 2);
        }
    return s;
}

int or_one_electron(int nshell, const int* l, const int* nprim, const int* prim_off, const double* exps,
                    const double* coefs, const double* centers, const int* bf_off, int nbf, int natom, const int* Z,
                    const double* xyz, double* S, double* T, double* V) {
    const double PI = 3.14159265358979323846;
    for (int i = 0; i < nbf * nbf; ++i) S[i] = T[i] = V[i] = 0.0;
    int ca_[15][3], cb_[15][3];
    for (int si = 0; si < nshell; ++si)
        for (int sj = 0; sj < nshell; ++sj) {
            const int la = l[si], lb = l[sj];
            cart_comps(la, ca_);
            cart_comps(lb, cb_);
            const double* A = centers + 3 * si;
            const double* B = centers + 3 * sj;
            double ab2 = 0.0;
            for (int x = 0; x < 3; ++x) ab2 += (A[x] - B[x]) * (A[x] - B[x]);
            for (int pa = 0; pa < nprim[si]; ++pa)
                for (int pb = 0; pb < nprim[sj]; ++pb) {
                    const double a = exps[prim_off[si] + pa], b = exps[prim_off[sj] + pb], p = a + b;
                    const double K = std::exp(-a * b / p * ab2);
                    double P[3];
                    for (int x = 0; x < 3; ++x) P[x] = (a * A[x] + b * B[x]) / p;
                    for (int ia = 0; ia < ncart(la); ++ia)
                        for (int jb = 0; jb < ncart(lb); ++jb) {
                            const int* ip = ca_[ia];
                            const int* jp = cb_[jb];
                            const double c = coefs[prim_off[si] + pa] * coefs[prim_off[sj] + pb] * cnorm(a, ip) *
                                             cnorm(b, jp);
                            double s1[3], t1[3];
                            for (int x = 0; x < 3; ++x) {
                                const double mA = P[x] - A[x], mB = P[x] - B[x], v = 0.5 / p;
                                const int i = ip[x], j = jp[x];
                                s1[x] = gauss_moment_1d(i, j, mA, mB, v);
                                t1[x] = -0.5 * (j * (j - 1) * gauss_moment_1d(i, j - 2, mA, mB, v) -
                                                2.0 * b * (2 * j + 1) * s1[x] +
                                                4.0 * b * b * gauss_moment_1d(i, j + 2, mA, mB, v));
                            }
                            const double pre = K * std::pow(PI / p, 1.5) * c;
                            const int idx = (bf_off[si] + ia) * nbf + bf_off[sj] + jb;
                            S[idx] += pre * s1[0] * s1[1] * s1[2];
                            T[idx] += pre * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                            const int nr = (la + lb) / 2 + 1;
                            for (int cn = 0; cn < natom; ++cn) {
                                const double* C = xyz + 3 * cn;
                                double pc2 = 0.0;
                                for (int x = 0; x < 3; ++x) pc2 += (P[x] - C[x]) * (P[x] - C[x]);
                                double rr[MAXR], ww[MAXR];
                                if (or_rys_roots(nr, p * pc2, rr, ww)) return 1;
                                double acc = 0.0;
                                for (int r = 0; r < nr; ++r) {
                                    double prod = ww[r];
                                    for (int x = 0; x < 3; ++x) {
                                        const double m = P[x] - rr[r] * (P[x] - C[x]);
                                        prod *= gauss_moment_1d(ip[x], jp[x], m - A[x], m - B[x], (1.0 - rr[r]) / (2.0 * p));
                                    }
                                    acc += prod;
                                }
                                V[idx] += -Z[cn] * 2.0 * PI / p * K * c * acc;
                            }
                        }
                }
        }
    return 0;
}
}  // extern "C"
