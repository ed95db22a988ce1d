// Host SCF driver interface.
#pragma once
#include "../../include/qcg.h"
#include "qcg_common.h"

#include <deque>
#include <vector>

namespace qcg {

using ScfConfig = qcg_scf_config;
using ScfResult = qcg_scf_result;

// One molecule's SCF state: init -> (Fock build, accept_fock, step)* until accept_fock returns true.
struct ScfRun {
    struct Diis {
        int max_size = 8;
        std::deque<std::vector<double>> f, e;
        std::vector<double> extrapolate() const;
    };
    int nbf = 0, nocc = 0, iter = 0;
    ScfConfig cfg{};
    double e_nuc = 0.0, e_tot = 0.0, last_dE = 0.0, last_rms = 0.0;
    bool done = false, converged = false, have_last = false;
    std::vector<double> S, H, X, D, Dold, F, F_last, C, eps;
    Diis diis;
    int init(int nbf, int nocc, const double* S, const double* H, double e_nuc, const ScfConfig& cfg,
             const double* D_init);
    bool accept_fock(const double* F, double e_elec);
    void step();
    void diagonalize(const std::vector<double>& F);
    void fill_result(ScfResult* res) const;
};

// Fock callback: D (row-major, symmetric) -> F and the electronic energy for that D.
typedef int (*FockFn)(void* user, const double* D, double* F, double* e_elec);

int scf_driver(int nbf, int nocc, const double* S, const double* H, double e_nuc, const ScfConfig& cfg,
               FockFn fock, void* user, const double* D_init, ScfResult* res, double* D_out, double* C_out,
               double* eps_out, double* F_out);

}  // namespace qcg
