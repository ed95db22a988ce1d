// A-priori bound on the energy error of a mixed-precision J/K build (qcg_error_bound) and the
// per-iteration SCF trace (qcg_scf_history).
//
// An error delta in the unique quartet (ab|cd), applied to its 8 index images, moves
// E2 = 1/2 Tr D (J - a/2 K) by delta * deg * [4 D_ab D_cd - a (D_ac D_bd + D_ad D_bc)], deg being the
// quartet's symmetry weight (1/2 per coinciding shell pair, as in the J/K kernels).  With
// |delta| <= eps32 Q_ab Q_cd for every Cartesian component (Schwarz, compute_schwarz_bounds,
// integrals/eri/mod.rs:1263-1290) and the components summed, the products of D elements sum to
// products of the shell-block sums S_xy = sum |D_ij| over i in x, j in y.
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.h"

namespace {

// Relative error allowed per FP32 quartet against its Schwarz product: 2^-16 (256 FP32 units in
// the last place).  The FP32 quartets measure below 2^-20 of Q_ab Q_cd (the test that builds single
// integrals through J in both precisions), so the bound keeps a factor of 16 over the worst seen.
constexpr double kEps32 = 0x1p-16;

}  // namespace

extern "C" {

int qcg_error_bound(qcg_context* c, const double* D, double hf_frac, double* dE_bound) {
    if (!c || !D || !dE_bound) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "qcg_error_bound: single molecules only");
    *dE_bound = 0.0;
    const bool mixed = c->precision == QCG_PREC_MIXED && c->fp32_cutoff > c->schwarz_thr;
    if (!mixed) return QCG_OK;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = prepare_jk(c);  // the FP32 quartets are exactly the plan's part-1 work items
    if (rc) return rc;
    const HostBasis& B = c->basis;
    const int ns = B.nshell, nbf = B.nbf;
    std::vector<double> S((size_t)ns * ns, 0.0);
    for (int x = 0; x < ns; ++x)
        for (int y = 0; y < ns; ++y) {
            double s = 0.0;
            for (int i = 0; i < ncart(B.l[x]); ++i)
                for (int j = 0; j < ncart(B.l[y]); ++j) s += std::fabs(D[(size_t)(B.bf_off[x] + i) * nbf + B.bf_off[y] + j]);
            S[(size_t)x * ns + y] = s;
        }
    const double a = hf_frac;
    double total = 0.0;
    int cls = 0;
    for (int cb = kNumPairClasses - 1; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++cls) {
            const int i0 = c->item_off[cls][1], n = c->item_cnt[cls][1];
            const PairRec* bra = c->pairs.data() + c->class_off[cb];
            const PairRec* ket = c->pairs.data() + c->class_off[ck];
            const bool diag = cb == ck;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
            for (int k = 0; k < n; ++k) {
                const int4 it = c->items[(size_t)i0 + k];
                // a bra cut into primitive ranges is listed once per range: count the quartets once
                if (((it.w >> 24) & kItemMaxPrim) != 0 && ((it.w >> 16) & kItemMaxPrim) != 0) continue;
                const PairRec& P = bra[it.x];
                const double dp = (P.sa == P.sb) ? 0.5 : 1.0;
                const double* Sa = &S[(size_t)P.sa * ns];
                const double* Sb = &S[(size_t)P.sb * ns];
                const double sab = Sa[P.sb];
                double s = 0.0;
                int g = it.y, r = it.z;
                for (int j = 0; j < item_kets(it); ++j, ++r) {
                    while (r >= c->segs[g].y) r -= c->segs[g++].y;
                    const int q = c->segs[g].x + r;
                    const PairRec& Q = ket[q];
                    const double deg = dp * ((Q.sa == Q.sb) ? 0.5 : 1.0) * ((diag && q == P.pos) ? 0.5 : 1.0);
                    const double dens = 4.0 * sab * S[(size_t)Q.sa * ns + Q.sb] +
                                        a * (Sa[Q.sa] * Sb[Q.sb] + Sa[Q.sb] * Sb[Q.sa]);
                    s += P.Q * Q.Q * deg * dens;
                }
                total += s;
            }
        }
    *dE_bound = kEps32 * total;
    return QCG_OK;
}

int qcg_scf_history(const qcg_context* c, qcg_scf_iteration* out, int cap, int* n) {
    if (!c || !n || cap < 0 || (cap > 0 && !out)) return QCG_INVALID_ARG;
    *n = (int)c->scf_history.size();
    const int m = std::min(cap, *n);
    std::copy(c->scf_history.begin(), c->scf_history.begin() + m, out);
    return QCG_OK;
}

}  // extern "C"
