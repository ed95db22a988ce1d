// Host-side basis description and one-electron integrals.
#pragma once
#include <vector>

#include "onee_pair.h"
#include "qcg_common.h"

namespace qcg {

// Flat shell list (bohr), as passed through qcg_set_basis.
struct HostBasis {
    int nshell = 0, nbf = 0, natom = 0;
    std::vector<int> l, nprim, prim_off, shell_atom, bf_off, Z;
    std::vector<double> exps, coefs, centers, xyz;
};

inline BasisView basis_view(const HostBasis& B) {
    return BasisView{B.nshell,           B.nbf,          B.natom,         B.l.data(),       B.nprim.data(),
                     B.prim_off.data(),  B.shell_atom.data(), B.bf_off.data(), B.Z.data(),  B.exps.data(),
                     B.coefs.data(),     B.centers.data(), B.xyz.data()};
}

double nuclear_repulsion(const HostBasis& B);
void nuclear_repulsion_gradient(const HostBasis& B, double* grad);

}  // namespace qcg
