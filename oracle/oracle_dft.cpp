// CPU ORACLE — test infrastructure only.  FP64 restatement of qc-core's B3LYP exchange-correlation
// path (paths relative to crates/qc-core/src):
//
//   or_becke_grid     dft/grid.rs:103-198 (+ mura_knowles_radial :250, sg1_prune :315,
//                     becke_partition :400), Lebedev tables = published data (dft/lebedev.rs)
//                     pinned: exact integration of even monomials (tests) + IQCP B3LYP energies
//   or_basis_on_grid  dft/ks_scf.rs:115-255 (orbital/grid.rs:174 screening, :501 angular_factor)
//   or_b3lyp          dft/functionals.rs:179, :324, :691, :1170, :1625      pinned: libxc vectors
//   or_vxc            dft/ks_scf.rs:936 (chi D), :965 screening (1e-10 per 512-point block),
//                     :1028 density, :1150-1296 blocked V_xc and the (V+V^T)/2 symmetrisation
//   or_basis_hessian_on_grid  scf/gradient.rs:1938-2050 (same screening as the values)
//   or_xc_gradient    scf/gradient.rs:1687-1910 (xc_gradient_analytical, GGA branch: chi D,
//                     rho, grad rho, sigma, B3LYP, grad_chi_D, the two accumulation parts)
//                     pinned: finite differences of E_xc with the grid held fixed (tests)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../iqcp_amd/csrc/lebedev_table.inc"  // data: Lebedev-Laikov points (x,y,z,w)

extern "C" {

static const double PI_ = 3.14159265358979323846;

// ------------------------------------------------------------------ grid
static const double SG1R[19] = {0.0,    1.0000, 0.5882, 3.0769, 2.0513, 1.5385, 1.2308, 1.0256, 0.8791, 0.7692,
                                0.6838, 4.0909, 3.1579, 2.5714, 2.1687, 1.8750, 1.6514, 1.4754, 1.3333};
static const double SG1A[3][4] = {{0.25, 0.50, 1.00, 4.50}, {0.1667, 0.50, 0.90, 3.50}, {0.10, 0.40, 0.80, 2.50}};

static const double* leb(int n) {
    for (int i = 0; i < 6; ++i)
        if (qcg_leb_orders[i] == n) return qcg_leb_xyzw + 4 * qcg_leb_off[i];
    return nullptr;
}

// Returns the number of points written (cap = capacity of the outputs), -1 on error.
int64_t or_becke_grid(int natom, const int* Z, const double* axyz, int n_radial, int quality, int pruning,
                      int64_t cap, double* xyz, double* w, int* atom_of) {
    const int max_ang = quality == 0 ? 302 : 590;
    int64_t out = 0;
    for (int a = 0; a < natom; ++a) {
        const int z = Z[a];
        const double R = (z == 3 || z == 4 || z == 11 || z == 12 || z == 19 || z == 20) ? 7.0 : 5.2;
        std::vector<double> cx, cy, cz, raw;
        for (int i = 0; i < n_radial; ++i) {
            double x = (i + 0.5) / n_radial, x3 = x * x * x, o = 1.0 - x3;
            double r = -R * std::log(o);
            double dr = R * 3.0 * x * x / (o * n_radial);
            int nang = max_ang;
            if (pruning) {
                double ra = (z < 19 ? SG1R[z] : 1.0) + 1e-200;
                int row = z <= 2 ? 0 : (z <= 10 ? 1 : 2);
                int place = 0;
                for (int k = 0; k < 4; ++k) place += (r / ra > SG1A[row][k]);
                const int L[5] = {6, 38, 86, 194, 86};
                nang = L[place];
            }
            const double* g = leb(nang);
            if (!g) return -1;
            double rf = r * r * dr;
            for (int k = 0; k < nang; ++k) {
                cx.push_back(axyz[3 * a] + r * g[4 * k]);
                cy.push_back(axyz[3 * a + 1] + r * g[4 * k + 1]);
                cz.push_back(axyz[3 * a + 2] + r * g[4 * k + 2]);
                raw.push_back(rf * g[4 * k + 3]);
            }
        }
        size_t np = raw.size();
        std::vector<double> pw(np, 1.0);
        if (natom > 1) {
#pragma omp parallel for
            for (long p = 0; p < (long)np; ++p) {
                std::vector<double> gd(natom), P(natom, 1.0);
                for (int i = 0; i < natom; ++i) {
                    double dx = cx[p] - axyz[3 * i], dy = cy[p] - axyz[3 * i + 1], dz = cz[p] - axyz[3 * i + 2];
                    gd[i] = std::sqrt(dx * dx + dy * dy + dz * dz);
                }
                for (int i = 0; i < natom; ++i)
                    for (int j = 0; j < i; ++j) {
                        double dx = axyz[3 * i] - axyz[3 * j], dy = axyz[3 * i + 1] - axyz[3 * j + 1],
                               dz = axyz[3 * i + 2] - axyz[3 * j + 2];
                        double rij = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (rij < 1e-15) continue;
                        double mu = (gd[i] - gd[j]) / rij;
                        double f1 = mu * (3.0 - mu * mu) * 0.5;
                        double f2 = f1 * (3.0 - f1 * f1) * 0.5;
                        double g3 = f2 * (3.0 - f2 * f2) * 0.5;
                        P[i] *= 0.5 * (1.0 - g3);
                        P[j] *= 0.5 * (1.0 + g3);
                    }
                double s = 0.0;
                for (int i = 0; i < natom; ++i) s += P[i];
                pw[p] = std::fabs(s) > 1e-30 ? P[a] / s : 1.0 / natom;
            }
        }
        for (size_t p = 0; p < np; ++p) {
            double wf = raw[p] * pw[p];
            if (std::fabs(wf) > 1e-30) {
                if (out >= cap) return -1;
                xyz[3 * out] = cx[p];
                xyz[3 * out + 1] = cy[p];
                xyz[3 * out + 2] = cz[p];
                w[out] = wf;
                atom_of[out] = a;
                ++out;
            }
        }
    }
    return out;
}

// ------------------------------------------------------------------ functional
static void slater(double r, double& e, double& v) {
    if (r <= 1e-20) { e = v = 0; return; }
    double t = std::cbrt(r);
    e = -0.7385587663820224 * t;
    v = -0.9847450218426965 * t;
}
static void vwn5(double rho, double& e, double& v) {
    const double A = 0.0310907, X0 = -0.10498, B = 3.72744, C = 12.9352, Q = 6.15199081975908, XX0 = 12.5549141492;
    if (rho <= 1e-20) { e = v = 0; return; }
    double rs = std::cbrt(3.0 / (4.0 * PI_ * rho)), x = std::sqrt(rs), xs = x * x;
    double X = xs + B * x + C, t2b = 2 * x + B, at = std::atan(Q / t2b);
    double t1 = std::log(xs / X), t2 = 2 * B / Q * at, xm = x - X0;
    double t3 = -B * X0 / XX0 * (std::log(xm * xm / X) + 2 * (B + 2 * X0) / Q * at);
    double ec = A * (t1 + t2 + t3);
    double den = t2b * t2b + Q * Q, dat = -2 * Q / den;
    double d1 = 2 / x - t2b / X, d2 = 2 * B / Q * dat;
    double d3 = -B * X0 / XX0 * (2 / xm - t2b / X + 2 * (B + 2 * X0) / Q * dat);
    double dedx = A * (d1 + d2 + d3);
    e = ec;
    v = ec - (rs / 3.0) * (dedx / (2 * x));
}
static void b88(double rho, double sigma, double& e, double& vr, double& vs) {
    const double beta = 0.0042, CX = 0.7385587663820224;
    if (rho <= 1e-20) { e = vr = vs = 0; return; }
    double ra = rho / 2.0, sa = (sigma > 0 ? sigma : 0.0) / 4.0;
    double ra13 = std::cbrt(ra), ra43 = ra * ra13, two13 = std::cbrt(2.0);
    double es = -two13 * CX * ra13;
    double g = std::sqrt(sa);
    if (g < 1e-30) { slater(rho, e, vr); vs = 0; return; }
    double x = g / ra43, x2 = x * x, ash = std::asinh(x);
    double den = 1.0 + 6.0 * beta * x * ash;
    double de = -beta * ra13 * x2 / den;
    e = es + de;
    double dden = 6.0 * beta * (ash + x * (1.0 / std::sqrt(x2 + 1.0)));
    double dg = (2.0 * x * den - x2 * dden) / (den * den);
    double dsl = -two13 * CX / (3.0 * ra13 * ra13);
    double dxr = -4.0 / 3.0 * x / ra;
    double dd = -beta * (x2 / den / (3.0 * ra13 * ra13) + ra13 * dg * dxr);
    vr = e + rho * (dsl + dd) * 0.5;
    double dxs = 1.0 / (2.0 * g * ra43);
    vs = rho * (-beta * ra13 * dg * dxs) * 0.25;
}
static void lyp(double rho, double sigma, double& e, double& vr, double& vs) {
    const double A = 0.04918, B = 0.132, C = 0.2533, D = 0.349, CF = 2.8712340001881918;
    if (rho <= 1e-20) { e = vr = vs = 0; return; }
    double s = sigma > 0 ? sigma : 0.0;
    double r13 = std::cbrt(rho), r23 = r13 * r13, rm13 = 1.0 / r13, r2 = rho * rho, r3 = r2 * rho;
    double rm113 = 1.0 / (r3 * r23), drm = D * rm13, inv = 1.0 / (1.0 + drm), crm = C * rm13;
    double om = std::exp(-crm) * inv * rm113, del = crm + drm * inv;
    double t1 = -A * rho * inv, r4 = r2 * r2, r143 = r4 * r23;
    double t2 = -A * B * CF * om * r143;
    double bc = A * B * om * r2 * (7.0 * del + 3.0) / 72.0;
    e = (t1 + t2 + bc * s) / rho;
    vs = bc;
    double h = 1.0 + drm;
    double dt1 = -A * (1.0 + 4.0 / 3.0 * drm) / (h * h);
    double dom = om * (rm13 / (3.0 * rho) * (C + D * inv) - 11.0 / (3.0 * rho));
    double r113 = r3 * r23;
    double dt2 = -A * B * CF * (dom * r143 + om * 14.0 / 3.0 * r113);
    double r43 = rho * r13;
    double ddel = -C * rm13 / (3.0 * rho) + D * inv / (3.0 * r43) * (-1.0 + drm * inv);
    double dq = ((dom * r2 + om * 2.0 * rho) * (7.0 * del + 3.0) + om * r2 * 7.0 * ddel) / 72.0;
    vr = dt1 + dt2 + A * B * dq * s;
}

int or_b3lyp(int64_t n, const double* rho, const double* sigma, double* exc, double* vrho, double* vsigma) {
    for (int64_t i = 0; i < n; ++i) {
        double es, vs_, ev, vv, eb, vrb, vsb, el, vrl, vsl;
        slater(rho[i], es, vs_);
        vwn5(rho[i], ev, vv);
        b88(rho[i], sigma[i], eb, vrb, vsb);
        lyp(rho[i], sigma[i], el, vrl, vsl);
        exc[i] = 0.08 * es + 0.72 * eb + 0.19 * ev + 0.81 * el;
        vrho[i] = 0.08 * vs_ + 0.72 * vrb + 0.19 * vv + 0.81 * vrl;
        vsigma[i] = 0.72 * vsb + 0.81 * vsl;
    }
    return 0;
}

// ------------------------------------------------------------------ basis on grid
static double cnorm_grid(double a, int i, int j, int k) {
    // cartesian_norm (orbital/grid.rs:543-561) == cartesian_gaussian_normalization
    double t = 2.0 * a / PI_;
    double base = std::sqrt(std::sqrt(t * t * t));
    int L = i + j + k;
    double fa = 4.0 * a;
    double num = L == 0 ? 1.0 : L == 1 ? std::sqrt(fa) : L == 2 ? fa : std::pow(fa, L / 2.0);
    auto df = [](int n) { double r = 1; for (int q = 2 * n - 1; q > 1; q -= 2) r *= q; return r; };
    return base * num / std::sqrt(df(i) * df(j) * df(k));
}
static double ipow(double x, int n) { double r = 1.0; for (int i = 0; i < n; ++i) r *= x; return r; }

// chi[g*nbf+mu], grad[g*nbf*3+mu*3+d] (row-major, the reference's layout)
int or_basis_on_grid(int nshell, const int* l, const int* nprim, const int* prim_off, const double* exps,
                     const double* coefs, const double* centers, const int* bf_off, int nbf, int64_t G,
                     const double* pts, double* chi, double* grad) {
    std::memset(chi, 0, sizeof(double) * G * nbf);
    std::memset(grad, 0, sizeof(double) * G * nbf * 3);
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < G; ++g)
        for (int s = 0; s < nshell; ++s) {
            double dx = pts[3 * g] - centers[3 * s], dy = pts[3 * g + 1] - centers[3 * s + 1],
                   dz = pts[3 * g + 2] - centers[3 * s + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            double amin = 1e300;
            for (int p = 0; p < nprim[s]; ++p) amin = std::fmin(amin, exps[prim_off[s] + p]);
            if (amin * r2 > 34.5388) continue;
            int n = 0;
            for (int i = l[s]; i >= 0; --i)
                for (int j = l[s] - i; j >= 0; --j, ++n) {
                    int k = l[s] - i - j;
                    int pw[3] = {i, j, k};
                    double d[3] = {dx, dy, dz};
                    double ang = ipow(dx, i) * ipow(dy, j) * ipow(dz, k);
                    double rad = 0.0;
                    for (int p = 0; p < nprim[s]; ++p) {
                        double a = exps[prim_off[s] + p];
                        rad += coefs[prim_off[s] + p] * cnorm_grid(a, i, j, k) * std::exp(-a * r2);
                    }
                    int mu = bf_off[s] + n;
                    if (std::fabs(ang) >= 1e-30) chi[g * nbf + mu] = ang * rad;
                    for (int dd = 0; dd < 3; ++dd) {
                        double dang = 0.0;
                        if (pw[dd] > 0) {
                            int q[3] = {i, j, k};
                            q[dd] -= 1;
                            dang = pw[dd] * ipow(dx, q[0]) * ipow(dy, q[1]) * ipow(dz, q[2]);
                        }
                        double gv = 0.0;
                        for (int p = 0; p < nprim[s]; ++p) {
                            double a = exps[prim_off[s] + p];
                            double nc = coefs[prim_off[s] + p] * cnorm_grid(a, i, j, k);
                            gv += nc * std::exp(-a * r2) * (dang - 2.0 * a * d[dd] * ang);
                        }
                        grad[g * nbf * 3 + mu * 3 + dd] = gv;
                    }
                }
        }
    return 0;
}

// ------------------------------------------------------------------ V_xc
// chi, grad from or_basis_on_grid; D symmetric nbf x nbf.  Returns Exc; V (nbf x nbf).
double or_vxc(int nbf, int64_t G, const double* chi, const double* grad, const double* w, const double* D,
              double* V, double* rho_out) {
    std::vector<double> X((size_t)G * nbf, 0.0);  // chi D
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < G; ++g)
        for (int k = 0; k < nbf; ++k) {
            double c = chi[g * nbf + k];
            if (c == 0.0) continue;
            for (int nu = 0; nu < nbf; ++nu) X[g * nbf + nu] += c * D[(size_t)k * nbf + nu];
        }
    std::vector<double> rho(G), sig(G), gr(3 * G), exc(G), vr(G), vs(G);
    for (int64_t g = 0; g < G; ++g) {
        double r = 0, gx = 0, gy = 0, gz = 0;
        for (int mu = 0; mu < nbf; ++mu) r += chi[g * nbf + mu] * X[g * nbf + mu];
        rho[g] = r > 0.0 ? r : 0.0;
        for (int mu = 0; mu < nbf; ++mu) {
            double cd = X[g * nbf + mu];
            gx += grad[g * nbf * 3 + mu * 3] * cd;
            gy += grad[g * nbf * 3 + mu * 3 + 1] * cd;
            gz += grad[g * nbf * 3 + mu * 3 + 2] * cd;
        }
        gr[3 * g] = 2 * gx;
        gr[3 * g + 1] = 2 * gy;
        gr[3 * g + 2] = 2 * gz;
        sig[g] = 4 * (gx * gx + gy * gy + gz * gz);
    }
    if (rho_out) std::memcpy(rho_out, rho.data(), G * sizeof(double));
    or_b3lyp(G, rho.data(), sig.data(), exc.data(), vr.data(), vs.data());
    double Exc = 0.0;
    for (int64_t g = 0; g < G; ++g) Exc += w[g] * exc[g] * rho[g];
    std::vector<double> Vm((size_t)nbf * nbf, 0.0);
    const int BLK = 512;  // VXC_GRID_BLOCK_SIZE (ks_scf.rs:1127)
    const bool screen = nbf > 30;  // screening used if nbf > 30 (ks_scf.rs:1666)
    for (int64_t g0 = 0; g0 < G; g0 += BLK) {
        int64_t g1 = g0 + BLK < G ? g0 + BLK : G;
        std::vector<int> sig_bf;
        for (int mu = 0; mu < nbf; ++mu) {
            double mx = 0.0;
            for (int64_t g = g0; g < g1; ++g) mx = std::fmax(mx, std::fabs(chi[g * nbf + mu]));
            if (!screen || mx > 1e-10) sig_bf.push_back(mu);  // GRID_SCREEN_THRESHOLD (:958)
        }
        if (sig_bf.empty()) continue;
        const bool use = screen && (int)sig_bf.size() * 4 < nbf * 3;  // :1201
        std::vector<int> cols;
        if (use) cols = sig_bf;
        else for (int mu = 0; mu < nbf; ++mu) cols.push_back(mu);
        const int nc = (int)cols.size();
        std::vector<double> aow((size_t)(g1 - g0) * nc);
        for (int64_t g = g0; g < g1; ++g) {
            double wv = w[g] * vr[g], f4 = 4.0 * w[g] * vs[g];
            for (int a = 0; a < nc; ++a) {
                int mu = cols[a];
                double dot = grad[g * nbf * 3 + mu * 3] * gr[3 * g] + grad[g * nbf * 3 + mu * 3 + 1] * gr[3 * g + 1] +
                             grad[g * nbf * 3 + mu * 3 + 2] * gr[3 * g + 2];
                aow[(g - g0) * nc + a] = wv * chi[g * nbf + mu] + f4 * dot;
            }
        }
#pragma omp parallel for schedule(static)
        for (int a = 0; a < nc; ++a)
            for (int b = 0; b < nc; ++b) {
                double s = 0.0;
                for (int64_t g = g0; g < g1; ++g) s += chi[g * nbf + cols[a]] * aow[(g - g0) * nc + b];
                Vm[(size_t)cols[a] * nbf + cols[b]] += s;
            }
    }
    for (int i = 0; i < nbf; ++i)
        for (int j = 0; j < nbf; ++j) V[(size_t)i * nbf + j] = 0.5 * (Vm[(size_t)i * nbf + j] + Vm[(size_t)j * nbf + i]);
    return Exc;
}

}  // extern "C"

extern "C" {
// hess[g*nbf*6 + mu*6 + idx], idx 0 xx, 1 xy, 2 xz, 3 yy, 4 yz, 5 zz (gradient.rs:1938-2050):
// d2 chi/dr_a dr_b = N e [d2ang - 2 alpha (r_b dang_a + r_a dang_b + delta_ab ang) + 4 alpha^2 r_a r_b ang]
int or_basis_hessian_on_grid(int nshell, const int* l, const int* nprim, const int* prim_off, const double* exps,
                             const double* coefs, const double* centers, const int* bf_off, int nbf, int64_t G,
                             const double* pts, double* hess) {
    std::memset(hess, 0, sizeof(double) * G * nbf * 6);
    static const int HA[6] = {0, 0, 0, 1, 1, 2}, HB[6] = {0, 1, 2, 1, 2, 2};
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < G; ++g)
        for (int s = 0; s < nshell; ++s) {
            const double d[3] = {pts[3 * g] - centers[3 * s], pts[3 * g + 1] - centers[3 * s + 1],
                                 pts[3 * g + 2] - centers[3 * s + 2]};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double amin = 1e300;
            for (int p = 0; p < nprim[s]; ++p) amin = std::fmin(amin, exps[prim_off[s] + p]);
            if (amin * r2 > 34.5388) continue;
            int n = 0;
            for (int i = l[s]; i >= 0; --i)
                for (int j = l[s] - i; j >= 0; --j, ++n) {
                    const int pw[3] = {i, j, l[s] - i - j};
                    auto angp = [&](int a0, int a1, int a2) {
                        if (a0 < 0 || a1 < 0 || a2 < 0) return 0.0;
                        return ipow(d[0], a0) * ipow(d[1], a1) * ipow(d[2], a2);
                    };
                    const double ang = angp(pw[0], pw[1], pw[2]);
                    double dang[3];
                    for (int a = 0; a < 3; ++a) {
                        int q[3] = {pw[0], pw[1], pw[2]};
                        q[a] -= 1;
                        dang[a] = pw[a] > 0 ? pw[a] * angp(q[0], q[1], q[2]) : 0.0;
                    }
                    const int mu = bf_off[s] + n;
                    for (int h = 0; h < 6; ++h) {
                        const int a = HA[h], b = HB[h];
                        double d2 = 0.0;
                        if (a == b) {
                            if (pw[a] >= 2) {
                                int q[3] = {pw[0], pw[1], pw[2]};
                                q[a] -= 2;
                                d2 = pw[a] * (pw[a] - 1) * angp(q[0], q[1], q[2]);
                            }
                        } else if (pw[a] > 0 && pw[b] > 0) {
                            int q[3] = {pw[0], pw[1], pw[2]};
                            q[a] -= 1;
                            q[b] -= 1;
                            d2 = pw[a] * pw[b] * angp(q[0], q[1], q[2]);
                        }
                        double v = 0.0;
                        for (int p = 0; p < nprim[s]; ++p) {
                            const double al = exps[prim_off[s] + p];
                            const double nc = coefs[prim_off[s] + p] * cnorm_grid(al, pw[0], pw[1], pw[2]);
                            const double e = std::exp(-al * r2);
                            v += nc * e *
                                 (d2 - 2.0 * al * (d[b] * dang[a] + d[a] * dang[b] + (a == b ? ang : 0.0)) +
                                  4.0 * al * al * d[a] * d[b] * ang);
                        }
                        hess[(g * nbf + mu) * 6 + h] = v;
                    }
                }
        }
    return 0;
}

// xc_gradient_analytical, GGA branch (scf/gradient.rs:1687-1882), B3LYP.  chi/grad/hess from the
// functions above; shell_atom/bf_off/l give each basis function's atom; grad (natom x 3) is ADDED to.
int or_xc_gradient(int nshell, const int* l, const int* bf_off, const int* shell_atom, int nbf, int64_t G,
                   const double* w, const double* chi, const double* gchi, const double* hess, const double* D,
                   double* grad) {
    std::vector<int> atom_of(nbf, 0);
    for (int s = 0; s < nshell; ++s)
        for (int c = 0; c < (l[s] + 1) * (l[s] + 2) / 2; ++c) atom_of[bf_off[s] + c] = shell_atom[s];
    int natom = 0;
    for (int s = 0; s < nshell; ++s) natom = std::max(natom, shell_atom[s] + 1);
    static const int HIDX[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    std::vector<double> acc((size_t)natom * 3, 0.0);
    for (int64_t g = 0; g < G; ++g) {
        const double* c = chi + g * nbf;
        const double* gc = gchi + g * nbf * 3;
        std::vector<double> X(nbf, 0.0), GX((size_t)nbf * 3, 0.0);  // chi D, grad_chi_D
        for (int mu = 0; mu < nbf; ++mu) {
            double x = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
            for (int nu = 0; nu < nbf; ++nu) {
                const double dmn = D[(size_t)mu * nbf + nu];
                x += c[nu] * dmn;
                gx += dmn * gc[nu * 3];
                gy += dmn * gc[nu * 3 + 1];
                gz += dmn * gc[nu * 3 + 2];
            }
            X[mu] = x;
            GX[mu * 3] = gx;
            GX[mu * 3 + 1] = gy;
            GX[mu * 3 + 2] = gz;
        }
        double r = 0.0, g3[3] = {0, 0, 0};
        for (int mu = 0; mu < nbf; ++mu) {
            r += c[mu] * X[mu];
            for (int d = 0; d < 3; ++d) g3[d] += gc[mu * 3 + d] * X[mu];
        }
        const double rho = std::fmax(r, 0.0);
        const double grho[3] = {2.0 * g3[0], 2.0 * g3[1], 2.0 * g3[2]};
        const double sigma = grho[0] * grho[0] + grho[1] * grho[1] + grho[2] * grho[2];
        double exc, vrho, vsig;
        or_b3lyp(1, &rho, &sigma, &exc, &vrho, &vsig);
        const double wvr = w[g] * vrho, wvs = w[g] * vsig;
        if (std::fabs(wvr) < 1e-30 && std::fabs(wvs) < 1e-30) continue;
        for (int mu = 0; mu < nbf; ++mu) {
            const int A = atom_of[mu];
            double aow = wvr * X[mu];
            if (std::fabs(wvs) > 1e-30)
                aow += 2.0 * wvs * (grho[0] * GX[mu * 3] + grho[1] * GX[mu * 3 + 1] + grho[2] * GX[mu * 3 + 2]);
            for (int d = 0; d < 3; ++d) acc[A * 3 + d] += -2.0 * aow * gc[mu * 3 + d];
            if (std::fabs(wvs) > 1e-30 && std::fabs(X[mu]) > 1e-30) {
                const double f2 = -4.0 * wvs * X[mu];
                const double* H = hess + (g * nbf + mu) * 6;
                for (int d = 0; d < 3; ++d) {
                    double hc = 0.0;
                    for (int e = 0; e < 3; ++e) hc += grho[e] * H[HIDX[d][e]];
                    acc[A * 3 + d] += f2 * hc;
                }
            }
        }
    }
    for (int i = 0; i < natom * 3; ++i) grad[i] += acc[i];
    return 0;
}

// Single functional component: 0 Slater, 1 VWN5, 2 B88, 3 LYP (vsigma = 0 for the LDA parts).
int or_xc_component(int which, int64_t n, const double* rho, const double* sigma, double* exc, double* vrho,
                    double* vsigma) {
    for (int64_t i = 0; i < n; ++i) {
        vsigma[i] = 0.0;
        switch (which) {
            case 0: slater(rho[i], exc[i], vrho[i]); break;
            case 1: vwn5(rho[i], exc[i], vrho[i]); break;
            case 2: b88(rho[i], sigma[i], exc[i], vrho[i], vsigma[i]); break;
            case 3: lyp(rho[i], sigma[i], exc[i], vrho[i], vsigma[i]); break;
            default: return 1;
        }
    }
    return 0;
}
}  // extern "C"
