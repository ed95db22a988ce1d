// Host FP64 SCF driver (the caller of the Fock build).
//
// Restates qc-core's rhf_scf (crates/qc-core/src/scf/mod.rs:1363-1586) and the iteration of
// ks_scf_with_guess (dft/ks_scf.rs:1874-2033): symmetric orthogonaliser X = S^-1/2
// (build_orthogonalizer :709, singular if min eigenvalue <= 1e-10), core-Hamiltonian guess,
// optional damping (:1430-1440), AO-basis DIIS on FDS - SDF (compute_diis_error :990,
// DiisState::extrapolate :1189 with the equal-weight fallback :1248-1286), level shift (:1015),
// convergence on |dE|, RMS dD and the orbital-gradient norm (:1514-1520).
// The Fock matrix comes from a callback so the same driver runs the RHF and KS builds.
#include "scf.h"

#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "linalg.h"

namespace qcg {

namespace {

struct Diis {
    int max_size, n2;
    std::deque<std::vector<double>> f, e;
    void push(const std::vector<double>& F, const std::vector<double>& E) {
        if ((int)f.size() >= max_size) {
            f.pop_front();
            e.pop_front();
        }
        f.push_back(F);
        e.push_back(E);
    }
    static double dot(const std::vector<double>& a, const std::vector<double>& b) {
        double s = 0.0;
        for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
        return s;
    }
    std::vector<double> average() const {
        std::vector<double> out(n2, 0.0);
        const double c = 1.0 / (double)f.size();
        for (auto& fi : f)
            for (int j = 0; j < n2; ++j) out[j] += c * fi[j];
        return out;
    }
    std::vector<double> extrapolate() const {
        const int n = (int)f.size();
        if (n == 1) return f[0];
        std::vector<double> B((size_t)(n + 1) * (n + 1), 0.0), rhs(n + 1, 0.0), c;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) B[(size_t)i * (n + 1) + j] = B[(size_t)j * (n + 1) + i] = dot(e[i], e[j]);
        for (int i = 0; i < n; ++i) B[(size_t)i * (n + 1) + n] = B[(size_t)n * (n + 1) + i] = -1.0;
        rhs[n] = -1.0;
        if (!lu_solve(B, rhs, n + 1, c)) return average();
        double cs = 0.0;
        for (int i = 0; i < n; ++i) cs += c[i];
        if (std::fabs(cs - 1.0) > 1e-6) return average();
        std::vector<double> out(n2, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n2; ++j) out[j] += c[i] * f[i][j];
        return out;
    }
};

}  // namespace

int scf_driver(int nbf, int nocc, const double* S, const double* H, double e_nuc, const ScfConfig& cfg,
               FockFn fock, void* user, const double* D_init, ScfResult* res, double* D_out, double* C_out,
               double* eps_out, double* F_out) {
    const int n = nbf;
    const size_t n2 = (size_t)n * n;
    if (n <= 0 || nocc < 0 || nocc > n) return QCG_INVALID_ARG;
    std::vector<double> Sv(S, S + n2), Hv(H, H + n2);
    // X = U diag(1/sqrt(l)) U^T
    std::vector<double> sw, sU;
    sym_eigen(Sv, n, sw, sU);
    if (sw[0] <= 1e-10) return QCG_NUMERICAL;  // ScfError::SingularOverlap
    std::vector<double> X(n2, 0.0);
    for (int k = 0; k < n; ++k) {
        const double f = 1.0 / std::sqrt(sw[k]);
        for (int i = 0; i < n; ++i) {
            const double a = sU[(size_t)i * n + k] * f;
            for (int j = 0; j < n; ++j) X[(size_t)i * n + j] += a * sU[(size_t)j * n + k];
        }
    }
    std::vector<double> tmp(n2), Fp(n2), C(n2), eps, Cp;
    auto diagonalize = [&](const std::vector<double>& F) {
        matmul(F.data(), X.data(), tmp.data(), n);          // F X
        matmul(X.data(), tmp.data(), Fp.data(), n, true);   // X^T F X
        sym_eigen(Fp, n, eps, Cp);
        matmul(X.data(), Cp.data(), C.data(), n);           // C = X C'
    };
    auto density = [&](std::vector<double>& D) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double s = 0.0;
                for (int k = 0; k < nocc; ++k) s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
                D[(size_t)i * n + j] = 2.0 * s;
            }
    };
    std::vector<double> D(n2), F(n2);
    if (D_init) {
        std::memcpy(D.data(), D_init, n2 * sizeof(double));
    } else {
        diagonalize(Hv);
        density(D);
    }
    double e_elec = 0.0;
    int rc = fock(user, D.data(), F.data(), &e_elec);
    if (rc) return rc;
    double e_tot = e_elec + e_nuc;
    Diis diis{cfg.diis_size > 0 ? cfg.diis_size : 1, (int)n2, {}, {}};
    std::vector<double> F_last;
    bool have_last = false, converged = false;
    int iter = 0;
    std::vector<double> FD(n2), FDS(n2), SD(n2), SDF(n2), Err(n2), Dold(n2);
    for (iter = 1; iter <= cfg.max_iterations; ++iter) {
        const double e_old = e_tot;
        Dold = D;
        std::vector<double> Fd = F;
        const bool damp = std::fabs(cfg.damp) > 1e-4 && have_last && (cfg.damp_start == 0 || iter < cfg.damp_start);
        if (damp)
            for (size_t i = 0; i < n2; ++i) Fd[i] = F_last[i] * cfg.damp + F[i] * (1.0 - cfg.damp);
        std::vector<double> Fdiag = Fd;
        if (cfg.use_diis && iter >= cfg.diis_start) {
            matmul(Fd.data(), D.data(), FD.data(), n);
            matmul(FD.data(), Sv.data(), FDS.data(), n);
            matmul(Sv.data(), D.data(), SD.data(), n);
            matmul(SD.data(), Fd.data(), SDF.data(), n);
            for (size_t i = 0; i < n2; ++i) Err[i] = FDS[i] - SDF[i];
            diis.push(Fd, Err);
            if (diis.f.size() >= 2) Fdiag = diis.extrapolate();
        }
        if (std::fabs(cfg.level_shift) >= 1e-10) {
            // F + shift (S - S (D/2) S)
            matmul(Sv.data(), D.data(), SD.data(), n);
            matmul(SD.data(), Sv.data(), tmp.data(), n);
            for (size_t i = 0; i < n2; ++i) Fdiag[i] += cfg.level_shift * (Sv[i] - 0.5 * tmp[i]);
        }
        diagonalize(Fdiag);
        density(D);
        rc = fock(user, D.data(), F.data(), &e_elec);
        if (rc) return rc;
        F_last = F;
        have_last = true;
        e_tot = e_elec + e_nuc;
        const double dE = std::fabs(e_tot - e_old);
        double ss = 0.0;
        for (size_t i = 0; i < n2; ++i) ss += (D[i] - Dold[i]) * (D[i] - Dold[i]);
        const double rms = std::sqrt(ss / (double)n2);
        // orbital gradient: C^T F C, virtual-occupied block
        matmul(F.data(), C.data(), tmp.data(), n);
        double g2 = 0.0;
        const int nvir = n - nocc;
        for (int v = nocc; v < n; ++v)
            for (int o = 0; o < nocc; ++o) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += C[(size_t)k * n + v] * tmp[(size_t)k * n + o];
                g2 += s * s;
            }
        const double gnorm = nvir > 0 ? std::sqrt(4.0 * g2) / std::sqrt((double)nvir * nocc) : 0.0;
        if (res) {
            res->last_delta_e = dE;
            res->last_rms_d = rms;
        }
        if (dE < cfg.energy_tol && rms < cfg.density_tol && gnorm < std::sqrt(cfg.energy_tol)) {
            converged = true;
            break;
        }
    }
    if (converged && std::fabs(cfg.level_shift) > 1e-10) diagonalize(F);
    if (res) {
        res->converged = converged ? 1 : 0;
        res->iterations = converged ? iter : cfg.max_iterations;
        res->energy_total = e_tot;
        res->energy_electronic = e_tot - e_nuc;
        res->energy_nuclear = e_nuc;
    }
    if (D_out) std::memcpy(D_out, D.data(), n2 * sizeof(double));
    if (C_out) std::memcpy(C_out, C.data(), n2 * sizeof(double));
    if (eps_out) std::memcpy(eps_out, eps.data(), n * sizeof(double));
    if (F_out) std::memcpy(F_out, F.data(), n2 * sizeof(double));
    return converged ? QCG_OK : QCG_NOT_CONVERGED;
}

}  // namespace qcg
