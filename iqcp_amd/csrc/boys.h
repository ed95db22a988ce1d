// Boys function F_0..F_m(T) in FP64, usable from host and device code.
//
// Follows qc-core's boys_all_internal (crates/qc-core/src/boys/mod.rs:587-621):
//   T < 1e-14          -> F_i = 1/(2i+1)
//   T < turnover(m)    -> power series for F_m (terms until x <= eps/2 * e), then downward recurrence
//   otherwise          -> F_0 = sqrt(pi/4T) erf(sqrt T), then upward recurrence
// The regime is chosen by the turnover point of the highest order m (TURNOVER_POINT, boys/mod.rs:67-108).
#pragma once
#include "qcg_common.h"

namespace qcg {

QCG_HD double boys_turnover(int m) {
    // TURNOVER_POINT values, boys/mod.rs:67-108; t0 = 0.5 ((2m-1)!!/sqrt(2m-1))^(1/(m-1))
    constexpr double tp[40] = {
        0.0, 0.0, 0.866025403784, 1.295010032056, 1.705493613097, 2.106432965305,
        2.501471934009, 2.892473348218, 3.280525047072, 3.666320693281, 4.05033123037,
        4.432891808508, 4.814249856864, 5.194593501454, 5.574069276051, 5.952793645111,
        6.330860773135, 6.708347923415, 7.08531930745, 7.461828891625, 7.837922483937,
        8.213639312398, 8.589013237349, 8.964073695432, 9.338846443746, 9.713354153046,
        10.08761688545, 10.46165248270, 10.83547688448, 11.20910439128, 11.58254788331,
        11.95581900374, 12.32892831326, 12.70188542111, 13.07469909673, 13.44737736550,
        13.81992759110, 14.19235654675, 14.56467047710, 14.93687515212};
    return m < 40 ? tp[m] : 14.93687515212 + 0.372 * (double)(m - 39);
}

// Returns 0 on success, QCG_INVALID_ARG for m > 50 or T < 0 (BoysError::InvalidOrder / NegativeArgument).
QCG_HD int boys_all(int m, double t, double* f) {
    if (m < 0 || m > 50 || !(t >= 0.0)) return QCG_INVALID_ARG;
    if (t < 1e-14) {
        for (int i = 0; i <= m; ++i) f[i] = 1.0 / (2.0 * i + 1.0);
        return QCG_OK;
    }
    if (t < boys_turnover(m)) {
        const double b = m + 0.5;
        const double e = 0.5 * exp(-t);
        double x = e, s = e;
        const double tol = 1.1102230246251565e-16 * e;  // DBL_EPSILON/2 * e
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
        const double tt = sqrt(t);
        f[0] = 0.88622692545275801364908374167057259 / tt * erf(tt);
        const double e = exp(-t);
        const double b = 0.5 / t;
        for (int i = 1; i <= m; ++i) f[i] = b * ((2 * i - 1) * f[i - 1] - e);
    }
    return QCG_OK;
}

}  // namespace qcg
