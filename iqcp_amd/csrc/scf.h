// Host SCF driver interface.
#pragma once
#include "../../include/qcg.h"
#include "qcg_common.h"

namespace qcg {

using ScfConfig = qcg_scf_config;
using ScfResult = qcg_scf_result;

// Fock callback: D (row-major, symmetric) -> F and the electronic energy for that D.
typedef int (*FockFn)(void* user, const double* D, double* F, double* e_elec);

int scf_driver(int nbf, int nocc, const double* S, const double* H, double e_nuc, const ScfConfig& cfg,
               FockFn fock, void* user, const double* D_init, ScfResult* res, double* D_out, double* C_out,
               double* eps_out, double* F_out);

}  // namespace qcg
