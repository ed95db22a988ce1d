// Closed-shell B3LYP (the VWN5 variant, B3LYP5) at one point, FP64, host and device.
//
// Written from the published functionals, each derivative worked out for this file:
//   Slater (Dirac) exchange, unpolarised:  f = -Cx rho^(4/3), Cx = (3/4)(3/pi)^(1/3)
//   VWN5 correlation (Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200 (1980), eq. 4.4 with the
//     paramagnetic "5" parameters), in x = sqrt(rs):
//       ec(x) = A [ ln(x^2/X) + (2b/Q) atn - (b x0/X0) ( ln((x-x0)^2/X) + (2(b+2x0)/Q) atn ) ]
//     X = x^2 + b x + c, X0 = X(x0), Q = sqrt(4c - b^2), atn = atan(Q/(2x+b));  since
//     (2x+b)^2 + Q^2 = 4X, d atn/dx = -Q/(2X), so every term of dec/dx is a ratio over X;
//     v = ec - (x/6) dec/dx.
//   Becke 88 exchange (A. D. Becke, Phys. Rev. A 38, 3098 (1988)) per spin channel with
//     rho_s = rho/2, g_s = |grad rho_s|^2 = sigma/4, x_s = sqrt(g_s) rho_s^(-4/3):
//       f = 2 [ -Cs rho_s^(4/3) - beta g_s rho_s^(-4/3) / d(x_s) ],  d = 1 + 6 beta x asinh(x),
//     Cs = 2^(1/3) Cx; the gradient term written with g_s rho_s^(-4/3) = rho_s^(4/3) x_s^2 so the
//     sigma derivative needs no division by |grad rho| (it is finite at sigma = 0).
//   LYP correlation (Lee, Yang, Parr, Phys. Rev. B 37, 785 (1988)) in the closed form of
//     Miehlich, Savin, Stoll, Preuss (Chem. Phys. Lett. 157, 200 (1989)), closed shell:
//     with u = rho^(-1/3):
//       f = -a rho/(1+du) - a b CF W + a b (3 + 7 delta)/72 Y sigma
//     W = rho e^(-cu)/(1+du) (= omega rho^(14/3)), Y = e^(-cu) rho^(-5/3)/(1+du) (= omega rho^2),
//     delta = c u + d u/(1+du); logarithmic derivatives: dlnW/drho = (1 + delta/3)/rho,
//     dlnY/drho = (delta - 5)/(3 rho), ddelta/drho = -(u/(3 rho)) (c + d/(1+du)^2).
//   B3LYP5 = 0.08 Slater + 0.72 B88 + 0.19 VWN5 + 0.81 LYP (+ 0.20 exact exchange elsewhere), the
//   combination of qc-core's B3lyp::eval_xc_gga (dft/functionals.rs:1625).
// Densities at or below 1e-20 give zeros (the reference's DENSITY_THRESHOLD, functionals.rs:31).
// exc is the energy per electron; vrho = d(rho exc)/d rho, vsigma = d(rho exc)/d sigma.
#pragma once
#include "qcg_common.h"

namespace qcg {

constexpr double kXcRhoThreshold = 1e-20;
constexpr double kB3lypA = 0.20, kB3lypB = 0.72, kB3lypC = 0.81;

namespace xcf {
constexpr double kCx = 0.73855876638202240588;        // (3/4)(3/pi)^(1/3)
constexpr double kTwoThird = 1.25992104989487316477;  // 2^(1/3)
// VWN5 paramagnetic parameters and the two derived constants
constexpr double kVA = 0.0310907, kVx0 = -0.10498, kVb = 3.72744, kVc = 12.9352;
constexpr double kVX0 = kVx0 * kVx0 + kVb * kVx0 + kVc;
constexpr double kVQ = 6.15199081975908023217;  // sqrt(4c - b^2)
// Becke 88
constexpr double kBeta = 0.0042;
// LYP
constexpr double kLa = 0.04918, kLb = 0.132, kLc = 0.2533, kLd = 0.349;
constexpr double kCF = 2.87123400018819181594;  // (3/10)(3 pi^2)^(2/3)
}  // namespace xcf

QCG_HD void xc_slater(double rho, double& e, double& v) {
    if (rho <= kXcRhoThreshold) { e = v = 0.0; return; }
    e = -xcf::kCx * cbrt(rho);
    v = (4.0 / 3.0) * e;
}

QCG_HD void xc_vwn5(double rho, double& e, double& v) {
    using namespace xcf;
    if (rho <= kXcRhoThreshold) { e = v = 0.0; return; }
    const double rs = cbrt(3.0 / (4.0 * kPi * rho));
    const double x = sqrt(rs);
    const double X = x * x + kVb * x + kVc;
    const double iX = 1.0 / X;
    const double atn = atan(kVQ / (2.0 * x + kVb));
    const double xm = x - kVx0;
    const double k0 = kVb * kVx0 / kVX0;  // weight of the x0 bracket
    const double k1 = 2.0 * kVb / kVQ, k2 = 2.0 * (kVb + 2.0 * kVx0) / kVQ;
    e = kVA * (log(x * x * iX) + k1 * atn - k0 * (log(xm * xm * iX) + k2 * atn));
    // d/dx: ln(x^2/X) -> 2/x - (2x+b)/X; (2b/Q) atn -> -b/X; ln((x-x0)^2/X) -> 2/(x-x0) - (2x+b)/X;
    // (2(b+2x0)/Q) atn -> -(b+2x0)/X
    const double dX = (2.0 * x + kVb) * iX;
    const double de = kVA * ((2.0 / x - dX - kVb * iX) - k0 * (2.0 / xm - dX - (kVb + 2.0 * kVx0) * iX));
    v = e - (x / 6.0) * de;
}

QCG_HD void xc_b88(double rho, double sigma, double& e, double& vr, double& vs) {
    using namespace xcf;
    if (rho <= kXcRhoThreshold) { e = vr = vs = 0.0; return; }
    const double rs = 0.5 * rho;                       // rho_s
    const double gs = 0.25 * (sigma > 0.0 ? sigma : 0.0);  // |grad rho_s|^2
    const double r13 = cbrt(rs);
    const double r43 = rs * r13;
    const double im43 = 1.0 / r43;                     // rho_s^(-4/3)
    const double x = sqrt(gs) * im43;
    const double ash = asinh(x);
    const double d = 1.0 + 6.0 * kBeta * x * ash;
    const double dp = 6.0 * kBeta * (ash + x / sqrt(1.0 + x * x));  // d'(x)
    const double id = 1.0 / d;
    const double cs = kTwoThird * kCx;
    const double grad_term = kBeta * gs * im43 * id;   // beta g_s rho_s^(-4/3) / d
    // f = 2 (-cs rho_s^(4/3) - grad_term); exc = f / rho = f / (2 rho_s)
    e = (-cs * r43 - grad_term) / rs;
    // vrho = d f / d rho = d g / d rho_s with g = -cs rho_s^(4/3) - grad_term;
    // d grad_term / d rho_s = -(4/3) grad_term (1 - x d'/d) / rho_s
    vr = -(4.0 / 3.0) * cs * r13 + (4.0 / 3.0) * grad_term * (1.0 - x * dp * id) / rs;
    // vsigma = d f / d sigma = (1/2) d g / d g_s;  d grad_term / d g_s = beta rho_s^(-4/3) (1 - x d'/(2d)) / d
    vs = -0.5 * kBeta * im43 * id * (1.0 - 0.5 * x * dp * id);
}

QCG_HD void xc_lyp(double rho, double sigma, double& e, double& vr, double& vs) {
    using namespace xcf;
    if (rho <= kXcRhoThreshold) { e = vr = vs = 0.0; return; }
    const double s = sigma > 0.0 ? sigma : 0.0;
    const double u = 1.0 / cbrt(rho);          // rho^(-1/3)
    const double h = 1.0 + kLd * u;
    const double ih = 1.0 / h;
    const double ex = exp(-kLc * u);
    const double delta = kLc * u + kLd * u * ih;
    const double ir = 1.0 / rho;
    const double W = rho * ex * ih;            // omega rho^(14/3)
    const double Y = ex * ih * u * u * u * u * u;  // omega rho^2 = e^(-cu) rho^(-5/3) / h
    const double f1 = -kLa * rho * ih;
    const double f2 = -kLa * kLb * kCF * W;
    const double gq = kLa * kLb * (3.0 + 7.0 * delta) / 72.0;  // sigma coefficient / Y
    const double f3 = gq * Y * s;
    e = (f1 + f2 + f3) * ir;
    vs = gq * Y;
    // d f1 / d rho = -a (1 + (4/3) d u) / h^2
    const double df1 = -kLa * (1.0 + (4.0 / 3.0) * kLd * u) * ih * ih;
    const double df2 = f2 * (1.0 + delta / 3.0) * ir;
    const double ddelta = -(u * ir / 3.0) * (kLc + kLd * ih * ih);
    const double df3 = s * (gq * Y * (delta - 5.0) * ir / 3.0 + kLa * kLb * (7.0 / 72.0) * Y * ddelta);
    vr = df1 + df2 + df3;
}

QCG_HD void xc_b3lyp(double rho, double sigma, double& exc, double& vrho, double& vsigma) {
    double es, vsl, ev, vv, eb, vrb, vsb, el, vrl, vsl2;
    xc_slater(rho, es, vsl);
    xc_vwn5(rho, ev, vv);
    xc_b88(rho, sigma, eb, vrb, vsb);
    xc_lyp(rho, sigma, el, vrl, vsl2);
    const double cs = 1.0 - kB3lypA - kB3lypB, cv = 1.0 - kB3lypC;
    exc = cs * es + kB3lypB * eb + cv * ev + kB3lypC * el;
    vrho = cs * vsl + kB3lypB * vrb + cv * vv + kB3lypC * vrl;
    vsigma = kB3lypB * vsb + kB3lypC * vsl2;
}

}  // namespace qcg
