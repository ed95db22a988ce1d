// Host-side launch interface of the stored-tensor Fock kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace qcg {

// F = H + G(D) from the compressed 8-fold tensor eri (npair(npair+1)/2 doubles, npair =
// nbf(nbf+1)/2, eri_index order); all pointers on the device, H, D, F nbf x nbf.
hipError_t launch_fock_stored(int nbf, const double* H, const double* D, const double* eri, double* F,
                              hipStream_t st);

}  // namespace qcg
