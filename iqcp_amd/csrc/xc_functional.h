// Closed-shell B3LYP (B3LYP5 = VWN5 variant) pointwise exchange-correlation, FP64, host/device.
//
// Same formulas as qc-core's dft/functionals.rs: SlaterExchange::eval (:179), Vwn5Correlation::eval
// (:324), Becke88Exchange::eval_gga (:691, spin-scaled, includes Slater), LypCorrelation::eval_gga
// (:1170), B3lyp::eval_xc_gga (:1625+) with a = 0.20, b = 0.72, c = 0.81:
//   exc   = 0.08 e_Slater + 0.72 e_B88 + 0.19 e_VWN5 + 0.81 e_LYP
//   vrho  = same weights on the potentials, vsigma = 0.72 vs_B88 + 0.81 vs_LYP
// rho <= 1e-20 gives zeros (DENSITY_THRESHOLD, :31).  exc is energy per electron.
#pragma once
#include "qcg_common.h"

namespace qcg {

constexpr double kXcRhoThreshold = 1e-20;
constexpr double kSlaterCx = 0.7385587663820224;        // (3/4)(3/pi)^(1/3)
constexpr double kSlater43Cx = 0.9847450218426965;      // (4/3) C_x
constexpr double kB3lypA = 0.20, kB3lypB = 0.72, kB3lypC = 0.81;

QCG_HD void xc_slater(double rho, double& e, double& v) {
    if (rho <= kXcRhoThreshold) { e = v = 0.0; return; }
    const double r3 = cbrt(rho);
    e = -kSlaterCx * r3;
    v = -kSlater43Cx * r3;
}

QCG_HD void xc_vwn5(double rho, double& e, double& v) {
    constexpr double A = 0.0310907, X0 = -0.10498, B = 3.72744, C = 12.9352;
    constexpr double Q = 6.15199081975908, XX0 = 12.5549141492;
    if (rho <= kXcRhoThreshold) { e = v = 0.0; return; }
    const double rs = cbrt(3.0 / (4.0 * kPi * rho));
    const double x = sqrt(rs);
    const double xs = x * x;
    const double X = xs + B * x + C;
    const double t2b = 2.0 * x + B;
    const double at = atan(Q / t2b);
    const double xm = x - X0;
    const double ec = A * (log(xs / X) + 2.0 * B / Q * at -
                           B * X0 / XX0 * (log(xm * xm / X) + 2.0 * (B + 2.0 * X0) / Q * at));
    const double den = t2b * t2b + Q * Q;
    const double dat = -2.0 * Q / den;
    const double dedx = A * ((2.0 / x - t2b / X) + 2.0 * B / Q * dat -
                             B * X0 / XX0 * (2.0 / xm - t2b / X + 2.0 * (B + 2.0 * X0) / Q * dat));
    e = ec;
    v = ec - (rs / 3.0) * (dedx / (2.0 * x));
}

QCG_HD void xc_b88(double rho, double sigma, double& e, double& vr, double& vs) {
    constexpr double beta = 0.0042;
    if (rho <= kXcRhoThreshold) { e = vr = vs = 0.0; return; }
    const double ra = 0.5 * rho;
    const double sa = fmax(sigma, 0.0) * 0.25;
    const double ra13 = cbrt(ra);
    const double ra43 = ra * ra13;
    const double two13 = 1.2599210498948731648;
    const double es_a = -two13 * kSlaterCx * ra13;
    const double g = sqrt(sa);
    if (g < 1e-30) {
        xc_slater(rho, e, vr);
        vs = 0.0;
        return;
    }
    const double x = g / ra43;
    const double x2 = x * x;
    const double ash = asinh(x);
    const double den = 1.0 + 6.0 * beta * x * ash;
    const double de_a = -beta * ra13 * x2 / den;
    e = es_a + de_a;
    const double dden = 6.0 * beta * (ash + x / sqrt(x2 + 1.0));
    const double dgdx = (2.0 * x * den - x2 * dden) / (den * den);
    const double des = -two13 * kSlaterCx / (3.0 * ra13 * ra13);
    const double dxdr = -4.0 / 3.0 * x / ra;
    const double dde = -beta * (x2 / den / (3.0 * ra13 * ra13) + ra13 * dgdx * dxdr);
    vr = e + rho * (des + dde) * 0.5;
    const double dxds = 1.0 / (2.0 * g * ra43);
    vs = rho * (-beta * ra13 * dgdx * dxds) * 0.25;
}

QCG_HD void xc_lyp(double rho, double sigma, double& e, double& vr, double& vs) {
    constexpr double A = 0.04918, B = 0.132, C = 0.2533, D = 0.349, CF = 2.8712340001881918;
    if (rho <= kXcRhoThreshold) { e = vr = vs = 0.0; return; }
    const double s = fmax(sigma, 0.0);
    const double r13 = cbrt(rho), r23 = r13 * r13, rm13 = 1.0 / r13;
    const double r2 = rho * rho, r3 = r2 * rho;
    const double rm113 = 1.0 / (r3 * r23);
    const double drm = D * rm13;
    const double inv = 1.0 / (1.0 + drm);
    const double crm = C * rm13;
    const double om = exp(-crm) * inv * rm113;
    const double del = crm + drm * inv;
    const double t1 = -A * rho * inv;
    const double r143 = r2 * r2 * r23;
    const double t2 = -A * B * CF * om * r143;
    const double bc = A * B * om * r2 * (7.0 * del + 3.0) / 72.0;
    e = (t1 + t2 + bc * s) / rho;
    vs = bc;
    const double h = 1.0 + drm;
    const double dt1 = -A * (1.0 + 4.0 / 3.0 * drm) / (h * h);
    const double dom = om * (rm13 / (3.0 * rho) * (C + D * inv) - 11.0 / (3.0 * rho));
    const double r113 = r3 * r23;
    const double dt2 = -A * B * CF * (dom * r143 + om * 14.0 / 3.0 * r113);
    const double r43 = rho * r13;
    const double ddel = -C * rm13 / (3.0 * rho) + D * inv / (3.0 * r43) * (-1.0 + drm * inv);
    const double dq = ((dom * r2 + om * 2.0 * rho) * (7.0 * del + 3.0) + om * r2 * 7.0 * ddel) / 72.0;
    vr = dt1 + dt2 + A * B * dq * s;
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
