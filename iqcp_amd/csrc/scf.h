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
    double e_nuc = 0.0, e_tot = 0.0, last_dE = 0.0, last_rms = 0.0, last_g = 0.0;
    bool done = false, converged = false, have_last = false;
    std::vector<double> S, H, X, D, Dold, F, F_last, C, eps;
    Diis diis;
    bool last_diis = false;  // the current D came from a DIIS-extrapolated Fock matrix
    std::vector<qcg_scf_iteration> history;  // ScfOutput.trace (scf/mod.rs:1407-1531)
    int init(int nbf, int nocc, const double* S, const double* H, double e_nuc, const ScfConfig& cfg,
             const double* D_init);
    // may_converge = false: the build was reduced precision, so convergence is not declared;
    // bound: that build's energy error bound (recorded in the trace)
    bool accept_fock(const double* F, double e_elec, bool may_converge = true, double bound = 0.0);
    bool criteria_met() const;
    void step();
    void diagonalize(const std::vector<double>& F);
    void fill_result(ScfResult* res) const;
};

// Fock callback: D (row-major, symmetric) -> F and the electronic energy for that D.
typedef int (*FockFn)(void* user, const double* D, double* F, double* e_elec);
// Switches the Fock callback from reduced to full (FP64) precision for the rest of the run.
typedef int (*RefineFn)(void* user);
// Energy error bound of a reduced-precision Fock build at D (recorded in the iteration trace).
typedef int (*BoundFn)(void* user, const double* D, double* bound);

// Density change below which a reduced-precision run switches to FP64 builds (mixed precision:
// the FP32 parts leave noise of about 1e-7 in F, so the tight criteria are only tested on FP64
// builds; the converged state and energy are those of the FP64 operator).
constexpr double kRefineRms = 1e-5;

// refine != nullptr: the callback starts in reduced precision; the driver calls refine once, when
// the density change drops below kRefineRms or the criteria are met, and declares convergence
// only on FP64 builds.
int scf_driver(int nbf, int nocc, const double* S, const double* H, double e_nuc, const ScfConfig& cfg,
               FockFn fock, void* user, const double* D_init, ScfResult* res, double* D_out, double* C_out,
               double* eps_out, double* F_out, RefineFn refine = nullptr, BoundFn bound = nullptr,
               std::vector<qcg_scf_iteration>* history = nullptr);

}  // namespace qcg
