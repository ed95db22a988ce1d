// Small dense FP64 linear algebra for the host SCF driver (row-major, n x n).
//   sym_eigen : Householder tridiagonalisation + implicit symmetric QR with Wilkinson shifts
//               (Golub & Van Loan 8.3); eigenvalues ascending, eigenvectors in columns (the role
//               of nalgebra SymmetricEigen + sorted_eigen, scf/mod.rs:743-778).
//   lu_solve  : Gaussian elimination with partial pivoting; returns false on a zero pivot
//               (nalgebra's LU solve returning None, used by DIIS, scf/mod.rs:1189-1262).
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>
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

// A (symmetric) -> w (ascending), V (columns = eigenvectors), row-major.
// Golub & Van Loan, Matrix Computations (4th ed.), section 8.3: Householder reduction to
// tridiagonal form (Algorithm 8.3.1) with the reflections accumulated into Q, then the implicit
// symmetric QR step with a Wilkinson shift (Algorithm 8.3.2) on the unreduced trailing block
// until every off-diagonal element is negligible.  The transformation is kept transposed (QT,
// one eigenvector per row) so that every reflection and Givens rotation updates contiguous rows.
inline void sym_eigen(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    std::vector<double> QT((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) QT[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), q(n), u(n);
    // ---- tridiagonalisation: reflection k zeroes A[k+2.., k] and A[k, k+2..]
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;  // length of x = A[k+1.., k]
        double xn2 = 0.0;
        for (int i = 1; i < m; ++i) xn2 += a(k + 1 + i, k) * a(k + 1 + i, k);
        if (xn2 == 0.0) continue;  // already tridiagonal in this column
        const double x0 = a(k + 1, k);
        const double mu = std::sqrt(x0 * x0 + xn2);
        // v = x - alpha e1 with alpha = -sign(x0) |x| (no cancellation); P = I - beta v v^T
        const double alpha = x0 >= 0.0 ? -mu : mu;
        v[0] = x0 - alpha;
        for (int i = 1; i < m; ++i) v[i] = a(k + 1 + i, k);
        const double beta = 1.0 / (mu * (mu + std::fabs(x0)));  // 2 / (v^T v)
        // trailing block B = A[k+1.., k+1..]: B <- P B P = B - v w^T - w v^T,
        // p = beta B v, w = p - (beta p^T v / 2) v
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            const double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
            for (int j = 0; j < m; ++j) s += row[j] * v[j];
            p[i] = beta * s;
        }
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * v[i];
        const double half = 0.5 * beta * pv;
        for (int i = 0; i < m; ++i) q[i] = p[i] - half * v[i];
        for (int i = 0; i < m; ++i) {
            double* row = &A[(size_t)(k + 1 + i) * n + k + 1];
            for (int j = 0; j < m; ++j) row[j] -= v[i] * q[j] + q[i] * v[j];
        }
        a(k + 1, k) = a(k, k + 1) = alpha;
        for (int i = 1; i < m; ++i) a(k + 1 + i, k) = a(k, k + 1 + i) = 0.0;
        // QT <- P QT on rows k+1.. (QT holds (P_k ... P_0) so that QT A0 QT^T = T):
        // u = beta v^T QT[k+1.., :], then QT[k+1+i, :] -= v_i u
        std::fill(u.begin(), u.end(), 0.0);
        for (int i = 0; i < m; ++i) {
            const double* row = &QT[(size_t)(k + 1 + i) * n];
            for (int j = 0; j < n; ++j) u[j] += v[i] * row[j];
        }
        for (int i = 0; i < m; ++i) {
            double* row = &QT[(size_t)(k + 1 + i) * n];
            const double f = beta * v[i];
            for (int j = 0; j < n; ++j) row[j] -= f * u[j];
        }
    }
    std::vector<double> d(n), e(n > 0 ? n : 1, 0.0);  // diagonal, e[i] couples i and i+1
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
    // ---- implicit symmetric QR with Wilkinson shifts
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1;
    int guard = 0;
    while (hi > 0 && guard < 64 * n) {
        // deflate negligible couplings at the bottom
        if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
            e[hi - 1] = 0.0;
            --hi;
            continue;
        }
        int lo = hi - 1;  // start of the unreduced block ending at hi
        while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) --lo;
        if (lo > 0) e[lo - 1] = 0.0;
        ++guard;
        // Wilkinson shift: the eigenvalue of the trailing 2x2 closer to d[hi]
        const double dd = 0.5 * (d[hi - 1] - d[hi]);
        const double eh = e[hi - 1];
        const double r = std::hypot(dd, eh);
        const double shift = d[hi] - eh * eh / (dd + (dd >= 0.0 ? r : -r));
        // bulge chase with Givens rotations on rows/columns (k, k+1), k = lo .. hi-1
        double x = d[lo] - shift, z = e[lo];
        for (int k = lo; k < hi; ++k) {
            const double h = std::hypot(x, z);
            const double c = h == 0.0 ? 1.0 : x / h, s = h == 0.0 ? 0.0 : z / h;
            if (k > lo) e[k - 1] = h;
            // T <- G^T T G on the 2x2 block (k, k+1) and the bulge
            const double dk = d[k], dk1 = d[k + 1], ek = e[k];
            d[k] = c * c * dk + 2.0 * c * s * ek + s * s * dk1;
            d[k + 1] = s * s * dk - 2.0 * c * s * ek + c * c * dk1;
            e[k] = c * s * (dk1 - dk) + (c * c - s * s) * ek;
            if (k + 1 < hi) {
                x = e[k];
                z = s * e[k + 1];
                e[k + 1] *= c;
            }
            double* r0 = &QT[(size_t)k * n];
            double* r1 = &QT[(size_t)(k + 1) * n];
            for (int j = 0; j < n; ++j) {
                const double u0 = r0[j], u1 = r1[j];
                r0[j] = c * u0 + s * u1;
                r1[j] = c * u1 - s * u0;
            }
        }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    w.resize(n);
    V.assign((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        w[c] = d[idx[c]];
        const double* row = &QT[(size_t)idx[c] * n];
        for (int r = 0; r < n; ++r) V[(size_t)r * n + c] = row[r];
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
