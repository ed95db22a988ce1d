// Small dense FP64 linear algebra for the host SCF driver (row-major, n x n).
//   sym_eigen : Householder tridiagonalisation + implicit QL with Wilkinson shifts; eigenvalues
//               ascending, eigenvectors in columns (the role of nalgebra SymmetricEigen +
//               sorted_eigen, scf/mod.rs:743-778).
//   lu_solve  : Gaussian elimination with partial pivoting; returns false on a zero pivot
//               (nalgebra's LU solve returning None, used by DIIS, scf/mod.rs:1189-1262).
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace qcg {

inline void matmul(const double* A, const double* B, double* C, int n, bool ta = false, bool tb = false) {
    std::vector<double> tmp((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double a = ta ? A[(size_t)k * n + i] : A[(size_t)i * n + k];
            if (a == 0.0) continue;
            double* row = &tmp[(size_t)i * n];
            if (!tb) {
                const double* b = &B[(size_t)k * n];
                for (int j = 0; j < n; ++j) row[j] += a * b[j];
            } else {
                for (int j = 0; j < n; ++j) row[j] += a * B[(size_t)j * n + k];
            }
        }
    std::copy(tmp.begin(), tmp.end(), C);
}

// A (symmetric, destroyed) -> w (ascending), V (columns = eigenvectors), row-major.
inline void sym_eigen(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
    // Householder reduction of A to tridiagonal form, accumulating Q in A (classic scheme).
    std::vector<double> d(n), e(n);
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    for (int i = n - 1; i > 0; --i) {
        const int l = i - 1;
        double h = 0.0;
        if (l > 0) {
            double scale = 0.0;
            for (int k = 0; k <= l; ++k) scale += std::fabs(a(i, k));
            if (scale == 0.0) {
                e[i] = a(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    a(i, k) /= scale;
                    h += a(i, k) * a(i, k);
                }
                double f = a(i, l);
                double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                a(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    a(j, i) = a(i, j) / h;
                    g = 0.0;
                    for (int k = 0; k <= j; ++k) g += a(j, k) * a(i, k);
                    for (int k = j + 1; k <= l; ++k) g += a(k, j) * a(i, k);
                    e[j] = g / h;
                    f += e[j] * a(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    f = a(i, j);
                    e[j] = g = e[j] - hh * f;
                    for (int k = 0; k <= j; ++k) a(j, k) -= (f * e[k] + g * a(i, k));
                }
            }
        } else {
            e[i] = a(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        const int l = i - 1;
        if (d[i] != 0.0) {
            for (int j = 0; j <= l; ++j) {
                double g = 0.0;
                for (int k = 0; k <= l; ++k) g += a(i, k) * a(k, j);
                for (int k = 0; k <= l; ++k) a(k, j) -= g * a(k, i);
            }
        }
        d[i] = a(i, i);
        a(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) a(j, i) = a(i, j) = 0.0;
    }
    // Implicit QL on the tridiagonal (d, e) with the transformation accumulated into A.
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-16 * dd) break;
            }
            if (m != l) {
                if (++iter > 60) break;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    e[i + 1] = (r = std::hypot(f, g));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = a(k, i + 1);
                        a(k, i + 1) = s * a(k, i) + c * f;
                        a(k, i) = c * a(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    V.assign((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        w[c] = d[idx[c]];
        for (int r = 0; r < n; ++r) V[(size_t)r * n + c] = a(r, idx[c]);
    }
}

// Solve M x = b (M n x n row-major).  Returns false on an exactly singular pivot.
inline bool lu_solve(std::vector<double> M, std::vector<double> b, int n, std::vector<double>& x) {
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(M[(size_t)i * n + k]) > std::fabs(M[(size_t)piv * n + k])) piv = i;
        if (M[(size_t)piv * n + k] == 0.0) return false;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(M[(size_t)k * n + j], M[(size_t)piv * n + j]);
            std::swap(b[k], b[piv]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = M[(size_t)i * n + k] / M[(size_t)k * n + k];
            if (f == 0.0) continue;
            for (int j = k; j < n; ++j) M[(size_t)i * n + j] -= f * M[(size_t)k * n + j];
            b[i] -= f * b[k];
        }
    }
    x.assign(n, 0.0);
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= M[(size_t)i * n + j] * x[j];
        x[i] = s / M[(size_t)i * n + i];
    }
    return true;
}

}  // namespace qcg
