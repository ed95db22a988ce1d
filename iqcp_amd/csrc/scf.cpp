// Host FP64 SCF driver (the caller of the Fock build).
//
// Restates qc-core's rhf_scf (crates/qc-core/src/scf/mod.rs:1363-1586) and the iteration of
// ks_scf_with_guess (dft/ks_scf.rs:1874-2033): symmetric orthogonaliser X = S^-1/2
// (build_orthogonalizer :709, singular if min eigenvalue <= 1e-10), core-Hamiltonian guess,
// optional damping (:1430-1440), AO-basis DIIS on FDS - SDF (compute_diis_error :990,
// DiisState::extrapolate :1189 with the equal-weight fallback :1248-1286), level shift (:1015),
// convergence on |dE|, RMS dD and the orbital-gradient norm (:1514-1520).
// ScfRun holds one molecule's state; scf_driver runs it against a Fock callback, and the batch
// driver (qcg_batch_rhf_scf) steps many of them in lockstep around one batched Fock build.
// A mixed-precision run starts on reduced-precision builds and finishes on FP64 ones (refine).
#include "scf.h"

#include <cmath>
#include <cstring>
#include <vector>

#include "linalg.h"

namespace qcg {

std::vector<double> ScfRun::Diis::extrapolate() const {
    const int n = (int)f.size();
    if (n == 1) return f[0];
    auto dot = [](const std::vector<double>& a, const std::vector<double>& b) {
        double s = 0.0;
        for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
        return s;
    };
    auto average = [&]() {
        std::vector<double> out(f[0].size(), 0.0);
        const double c = 1.0 / (double)n;
        for (auto& fi : f)
            for (size_t j = 0; j < out.size(); ++j) out[j] += c * fi[j];
        return out;
    };
    std::vector<double> B((size_t)(n + 1) * (n + 1), 0.0), rhs(n + 1, 0.0), c;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) B[(size_t)i * (n + 1) + j] = B[(size_t)j * (n + 1) + i] = dot(e[i], e[j]);
    for (int i = 0; i < n; ++i) B[(size_t)i * (n + 1) + n] = B[(size_t)n * (n + 1) + i] = -1.0;
    rhs[n] = -1.0;
    if (!lu_solve(B, rhs, n + 1, c)) return average();
    double cs = 0.0;
    for (int i = 0; i < n; ++i) cs += c[i];
    if (std::fabs(cs - 1.0) > 1e-6) return average();
    std::vector<double> out(f[0].size(), 0.0);
    for (int i = 0; i < n; ++i)
        for (size_t j = 0; j < out.size(); ++j) out[j] += c[i] * f[i][j];
    return out;
}

void ScfRun::diagonalize(const std::vector<double>& Fm) {
    const int n = nbf;
    std::vector<double> tmp((size_t)n * n), Fp((size_t)n * n), Cp;
    matmul(Fm.data(), X.data(), tmp.data(), n);         // F X
    matmul(X.data(), tmp.data(), Fp.data(), n, true);   // X^T F X
    sym_eigen(Fp, n, eps, Cp);
    C.assign((size_t)n * n, 0.0);
    matmul(X.data(), Cp.data(), C.data(), n);           // C = X C'
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < nocc; ++k) s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
            D[(size_t)i * n + j] = 2.0 * s;
        }
}

int ScfRun::init(int nbf_, int nocc_, const double* S_, const double* H_, double e_nuc_, const ScfConfig& cfg_,
                 const double* D_init) {
    nbf = nbf_;
    nocc = nocc_;
    cfg = cfg_;
    e_nuc = e_nuc_;
    const size_t n2 = (size_t)nbf * nbf;
    if (nbf <= 0 || nocc < 0 || nocc > nbf) return QCG_INVALID_ARG;
    S.assign(S_, S_ + n2);
    H.assign(H_, H_ + n2);
    std::vector<double> sw, sU;
    sym_eigen(S, nbf, sw, sU);
    if (sw[0] <= 1e-10) return QCG_NUMERICAL;  // ScfError::SingularOverlap
    X.assign(n2, 0.0);
    for (int k = 0; k < nbf; ++k) {
        const double f = 1.0 / std::sqrt(sw[k]);
        for (int i = 0; i < nbf; ++i) {
            const double a = sU[(size_t)i * nbf + k] * f;
            for (int j = 0; j < nbf; ++j) X[(size_t)i * nbf + j] += a * sU[(size_t)j * nbf + k];
        }
    }
    D.assign(n2, 0.0);
    if (D_init) std::memcpy(D.data(), D_init, n2 * sizeof(double));
    else diagonalize(H);
    diis.max_size = cfg.diis_size > 0 ? cfg.diis_size : 1;
    iter = 0;
    done = converged = false;
    have_last = false;
    return QCG_OK;
}

// F and E_elec for the current D.  Returns true when this molecule is finished.
bool ScfRun::criteria_met() const {
    return last_dE < cfg.energy_tol && last_rms < cfg.density_tol && last_g < std::sqrt(cfg.energy_tol);
}

bool ScfRun::accept_fock(const double* F_, double e_elec, bool may_converge, double bound) {
    const size_t n2 = (size_t)nbf * nbf;
    F.assign(F_, F_ + n2);
    const double e_new = e_elec + e_nuc;
    qcg_scf_iteration rec{};
    rec.iteration = iter;
    rec.diis_applied = iter > 0 && last_diis;
    rec.fp32_build = may_converge ? 0 : 1;
    rec.energy_total = e_new;
    rec.energy_electronic = e_elec;
    rec.delta_e = rec.rms_density_change = std::nan("");
    rec.error_bound = bound;
    struct Push {  // records the iteration on every return path below
        ScfRun* r;
        qcg_scf_iteration* rec;
        ~Push() {
            if (r->iter > 0) {
                rec->delta_e = r->last_dE;
                rec->rms_density_change = r->last_rms;
            }
            rec->converged = r->converged ? 1 : 0;
            r->history.push_back(*rec);
        }
    } push{this, &rec};
    if (iter > 0) {
        F_last = F;
        have_last = true;
        last_dE = std::fabs(e_new - e_tot);
        double ss = 0.0;
        for (size_t i = 0; i < n2; ++i) ss += (D[i] - Dold[i]) * (D[i] - Dold[i]);
        last_rms = std::sqrt(ss / (double)n2);
        std::vector<double> tmp(n2);
        matmul(F.data(), C.data(), tmp.data(), nbf);
        double g2 = 0.0;
        for (int v = nocc; v < nbf; ++v)
            for (int o = 0; o < nocc; ++o) {
                double s = 0.0;
                for (int k = 0; k < nbf; ++k) s += C[(size_t)k * nbf + v] * tmp[(size_t)k * nbf + o];
                g2 += s * s;
            }
        const int nvir = nbf - nocc;
        const double gnorm = nvir > 0 ? std::sqrt(4.0 * g2) / std::sqrt((double)nvir * nocc) : 0.0;
        e_tot = e_new;
        last_g = gnorm;
        if (may_converge && criteria_met()) {
            converged = done = true;
            if (std::fabs(cfg.level_shift) > 1e-10) diagonalize(F);
            return true;
        }
        if (iter >= cfg.max_iterations) {
            done = true;
            return true;
        }
    } else {
        e_tot = e_new;
    }
    return false;
}

// Next density from the current Fock matrix (damping, DIIS, level shift, diagonalisation).
// Damping mixes with the Fock matrix of the latest build, as rhf_scf does (fock_last is set right
// after every build, scf/mod.rs:1497-1500), so it only changes rounding.
void ScfRun::step() {
    const int n = nbf;
    const size_t n2 = (size_t)n * n;
    ++iter;
    Dold = D;
    std::vector<double> Fd = F;
    const bool damp = std::fabs(cfg.damp) > 1e-4 && have_last && (cfg.damp_start == 0 || iter < cfg.damp_start);
    if (damp)
        for (size_t i = 0; i < n2; ++i) Fd[i] = F_last[i] * cfg.damp + F[i] * (1.0 - cfg.damp);
    std::vector<double> Fdiag = Fd;
    // D S, shared by the DIIS error and the level shift; F, D and S are symmetric, so
    // S D F = (F D S)^T and the error FDS - SDF needs one product more
    std::vector<double> DS;
    const bool diis_now = cfg.use_diis && iter >= cfg.diis_start;
    const bool shift = std::fabs(cfg.level_shift) >= 1e-10;
    if (diis_now || shift) {
        DS.resize(n2);
        matmul(D.data(), S.data(), DS.data(), n);
    }
    if (diis_now) {
        std::vector<double> FDS(n2), Err(n2);
        matmul(Fd.data(), DS.data(), FDS.data(), n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Err[(size_t)i * n + j] = FDS[(size_t)i * n + j] - FDS[(size_t)j * n + i];
        if ((int)diis.f.size() >= diis.max_size) {
            diis.f.pop_front();
            diis.e.pop_front();
        }
        diis.f.push_back(Fd);
        diis.e.push_back(Err);
        if (diis.f.size() >= 2) Fdiag = diis.extrapolate();
    }
    last_diis = diis_now && diis.f.size() >= 2;
    if (shift) {
        std::vector<double> SDS(n2);
        matmul(S.data(), DS.data(), SDS.data(), n);
        for (size_t i = 0; i < n2; ++i) Fdiag[i] += cfg.level_shift * (S[i] - 0.5 * SDS[i]);
    }
    diagonalize(Fdiag);
}

void ScfRun::fill_result(ScfResult* res) const {
    if (!res) return;
    res->converged = converged ? 1 : 0;
    res->iterations = converged ? iter : cfg.max_iterations;
    res->energy_total = e_tot;
    res->energy_electronic = e_tot - e_nuc;
    res->energy_nuclear = e_nuc;
    res->last_delta_e = last_dE;
    res->last_rms_d = last_rms;
}

int scf_driver(int nbf, int nocc, const double* S, const double* H, double e_nuc, const ScfConfig& cfg,
               FockFn fock, void* user, const double* D_init, ScfResult* res, double* D_out, double* C_out,
               double* eps_out, double* F_out, RefineFn refine, BoundFn bound,
               std::vector<qcg_scf_iteration>* history) {
    ScfRun run;
    if (history) history->clear();
    int rc = run.init(nbf, nocc, S, H, e_nuc, cfg, D_init);
    if (rc) return rc;
    std::vector<double> F((size_t)nbf * nbf);
    bool exact = refine == nullptr;
    for (;;) {
        double e_elec = 0.0, b = 0.0;
        rc = fock(user, run.D.data(), F.data(), &e_elec);
        if (rc) return rc;
        if (!exact && bound) {
            rc = bound(user, run.D.data(), &b);
            if (rc) return rc;
        }
        const bool fin = run.accept_fock(F.data(), e_elec, exact, b);
        if (history) *history = run.history;
        if (fin) break;
        if (!exact && run.iter > 0 && (run.last_rms < kRefineRms || run.criteria_met())) {
            rc = refine(user);
            if (rc) return rc;
            exact = true;
        }
        run.step();
    }
    run.fill_result(res);
    const size_t n2 = (size_t)nbf * nbf;
    if (D_out) std::memcpy(D_out, run.D.data(), n2 * sizeof(double));
    if (C_out && !run.C.empty()) std::memcpy(C_out, run.C.data(), n2 * sizeof(double));
    if (eps_out && !run.eps.empty()) std::memcpy(eps_out, run.eps.data(), nbf * sizeof(double));
    if (F_out) std::memcpy(F_out, run.F.data(), n2 * sizeof(double));
    return run.converged ? QCG_OK : QCG_NOT_CONVERGED;
}

}  // namespace qcg
