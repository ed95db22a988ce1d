// Device-resident data layouts shared by the host library and the kernels.
#pragma once
#include "qcg_common.h"

namespace qcg {

// One primitive pair (i, j) of a shell pair, array-of-structs (48 B + pad) so that a lane
// reading "its" pair touches one or two cache lines.
//   p     = a_i + a_j
//   P     = (a_i A + a_j B) / p
//   kab   = exp(-a_i a_j / p |A-B|^2)   (raw, used by the 1e-15 primitive screens)
//   ck    = kab * c_i c_j * n(a_i, la) n(a_j, lb), n(a, l) = (2a/pi)^(3/4) (4a)^(l/2)
// The component-dependent part of the reference normalisation, 1/sqrt((2i-1)!!(2j-1)!!(2k-1)!!)
// (integrals/overlap.rs:343-377), is applied after contraction.
template <typename R>
struct alignas(8 * sizeof(R)) PrimPair {
    R p, px, py, pz, kab, ck, pad0, pad1;
};

// A shell pair with l(sa) >= l(sb) (the two shells are swapped if needed; integrals are
// invariant to that relabelling).  Lists of these are grouped by pair class and sorted by
// Schwarz bound Q, descending.
struct alignas(16) PairRec {
    double Q;           // sqrt(max |(ab|ab)|), compute_schwarz_bounds (integrals/eri/mod.rs:1263-1290)
    double ax, ay, az;  // centre of shell sa
    double abx, aby, abz;  // A - B
    int sa, sb;         // shell indices (sa may be > or < sb)
    int bfa, bfb;       // first basis function of each shell
    int pp_off, npp;    // primitive pairs [pp_off, pp_off + npp) surviving k_ab >= 1e-15
    int atom_a, atom_b;
    int mol;            // molecule of a batch (0 for a single molecule)
    int pos;            // position in its molecule's sorted class list
};

}  // namespace qcg
