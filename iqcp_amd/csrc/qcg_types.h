// Device-resident data layouts shared by the host library and the kernels.
#pragma once
#include "qcg_common.h"

namespace qcg {

// One primitive pair (i, j) of a shell pair, array-of-structs (8 values, one 64 B line in FP64) so
// that a lane reading "its" pair touches one cache line.
//   p     = a_i + a_j
//   P     = (a_i A + a_j B) / p
//   kab   = exp(-a_i a_j / p |A-B|^2)   (raw, used by the 1e-15 primitive screens)
//   ckp   = kab * c_i c_j * n(a_i, la) n(a_j, lb) / p, n(a, l) = (2a/pi)^(3/4) (4a)^(l/2), so that
//           the quartet prefactor 2 pi^(5/2) / (p q sqrt(p+q)) K_ab K_cd is 2 pi^(5/2) ckp_ab ckp_cd / sqrt(p+q)
//   hp    = 1 / (2p)
//   ea    = a_i (the gradient kernels take a_j as p - ea)
// The component-dependent part of the reference normalisation, 1/sqrt((2i-1)!!(2j-1)!!(2k-1)!!)
// (integrals/overlap.rs:343-377), is applied after contraction.
template <typename R>
struct alignas(8 * sizeof(R)) PrimPair {
    R p, px, py, pz, kab, ckp, hp, ea;
};

// k_ab of a primitive pair in FP64 for the 1e-15 primitive screens (k_ab k_cd < 1e-15 skips the
// primitive quartet, integrals/eri/mod.rs:611).  The FP32 copies carry the FP64 value's bits in their
// kab (low word) and ea (high word) fields, ea being read by the FP64 gradient kernels only, so the
// FP32 quartets take exactly the FP64 screening decisions (those the planner and the flop model
// count).
QCG_HD double prim_kab(const PrimPair<double>& a) { return a.kab; }
QCG_HD double prim_kab(const PrimPair<float>& a) {
    const unsigned long long lo = __builtin_bit_cast(unsigned int, a.kab), hi = __builtin_bit_cast(unsigned int, a.ea);
    return __builtin_bit_cast(double, lo | (hi << 32));
}
inline void set_prim_kab(PrimPair<float>& a, double kab) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, kab);
    a.kab = __builtin_bit_cast(float, (unsigned int)(b & 0xffffffffu));
    a.ea = __builtin_bit_cast(float, (unsigned int)(b >> 32));
}

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
