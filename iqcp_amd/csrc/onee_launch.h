// Host-side launch interface of the one-electron kernels.
#pragma once
#include "onee_pair.h"

namespace qcg {

// work[u] = (si, sj, pi, pj): one ordered shell pair and one of its primitive pairs per entry; the
// terms are ADDED into grad (natom x 3, device).  D, W: nbf x nbf on the device.
// work[u] = (si, sj, ia, jb), si <= sj: one Cartesian component pair per entry; S, T, V (nbf x nbf,
// device) receive every element of the listed shell pairs and of their mirror images.
hipError_t launch_onee_matrices(const BasisView& B, const int4* work, int nwork, const double* rys_tab, double* S,
                                double* T, double* V, hipStream_t st);

hipError_t launch_onee_grad(const BasisView& B, const int4* work, int nwork, const double* D, const double* W,
                            const double* rys_tab, double* grad, hipStream_t st);

}  // namespace qcg
