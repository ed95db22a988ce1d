// Host side of the analytic nuclear gradient: the adjoint ERI kernels on the device plus the
// one-electron and nuclear-repulsion terms on the host (rhf_gradient, scf/gradient.rs:900-976).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"

using namespace qcg;

namespace {

// Work items of the gradient kernels: every bra pair's surviving kets (the exact screening
// predicate Q_P Q_Q >= threshold, one segment per contraction bucket, as for J/K) cut into items
// of grad_item_kets(class pair) kets.

int eri_grad(qcg_context* c, const double* dD, double hf_frac, double* dgrad, hipStream_t st) {
    std::vector<int4> items;
    std::vector<int2> segs;
    int off[21], cnt[21];
    int cls = 0;
    for (int cb = kNumPairClasses - 1; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++cls) {
            off[cls] = (int)items.size();
            for (int p = 0; p < c->ket_cnt_h[cb]; ++p) {
                const int g0 = (int)segs.size();
                for_ket_ranges(c, cb, ck, 0, p, c->schwarz_thr, c->schwarz_thr,
                               [&](int bs, int, int e) { segs.push_back(make_int2(bs, e - bs)); });
                cut_items_split(items, segs, g0, p, grad_item_kets(cb, ck), c->pairs[c->class_off[cb] + p].npp,
                                c->prim_split);
            }
            cnt[cls] = (int)items.size() - off[cls];
        }
    if (items.empty()) return QCG_OK;
    int4* ditems = nullptr;
    int2* dsegs = nullptr;
    HIPCHK(c, c->alloc(&ditems, items.size()));
    HIPCHK(c, c->alloc(&dsegs, segs.size()));
    HIPCHK(c, hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(dsegs, segs.data(), segs.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    cls = 0;
    int rc = QCG_OK;
    for (int cb = kNumPairClasses - 1; cb >= 0 && rc == QCG_OK; --cb)
        for (int ck = cb; ck >= 0; --ck, ++cls) {
            GradLaunch a{};
            a.bra = c->d_pairs + c->class_off[cb];
            a.nbra = c->class_off[cb + 1] - c->class_off[cb];
            a.ket = c->d_pairs + c->class_off[ck];
            a.nket = c->class_off[ck + 1] - c->class_off[ck];
            a.diag = cb == ck;
            a.pp = c->d_pp;
            a.rys = c->d_rys;
            a.D = dD;
            a.nbf = c->basis.nbf;
            a.xfac = 0.25 * hf_frac;
            a.thr = c->schwarz_thr;
            a.natom = c->basis.natom;
            a.grad = dgrad;
            a.items = ditems + off[cls];
            a.segs = dsegs;
            a.nitems = cnt[cls];
            const hipError_t e = launch_grad(cb, ck, a, st);
            if (e != hipSuccess) {
                rc = fail(c, QCG_DEVICE, std::string("launch_grad: ") + hipGetErrorString(e));
                break;
            }
        }
    // the item buffer is released only after the kernels that read it have finished
    const hipError_t es = hipStreamSynchronize(st);
    c->release(ditems);
    c->release(dsegs);
    if (rc) return rc;
    HIPCHK(c, es);
    return QCG_OK;
}

}  // namespace

extern "C" {

int qcg_eri_gradient(qcg_context* c, const double* D, double hf_frac, double* grad) {
    if (!c || !D || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "gradient: one molecule per context");
    HIPCHK(c, hipSetDevice(c->device));
    const int nbf = c->basis.nbf, n3 = 3 * c->basis.natom;
    const size_t n2 = (size_t)nbf * nbf;
    double* dg = nullptr;
    HIPCHK(c, c->alloc(&dg, n3));
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_fill(dg, n3, 0.0, c->stream));
    int rc = eri_grad(c, c->d_D, hf_frac, dg, c->stream);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(grad, dg, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->release(dg);
    return QCG_OK;
}

int qcg_one_electron_gradient(qcg_context* c, const double* D, const double* W, double* grad) {
    if (!c || !D || !W || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    std::memset(grad, 0, 3 * c->basis.natom * sizeof(double));
    nuclear_repulsion_gradient(c->basis, grad);
    one_electron_gradient(c->basis, D, W, grad);
    return QCG_OK;
}

int qcg_rhf_gradient(qcg_context* c, const double* D, const double* C, const double* eps, int n_occ, double* grad) {
    if (!c || !D || !C || !eps || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    const int nbf = c->basis.nbf, n3 = 3 * c->basis.natom;
    if (n_occ < 0 || n_occ > nbf) return fail(c, QCG_INVALID_ARG, "n_occ");
    // W = 2 sum_i eps_i C_mu,i C_nu,i  (build_energy_weighted_density, scf/gradient.rs:101)
    std::vector<double> W((size_t)nbf * nbf, 0.0);
    for (int mu = 0; mu < nbf; ++mu)
        for (int nu = 0; nu < nbf; ++nu) {
            double v = 0.0;
            for (int i = 0; i < n_occ; ++i) v += eps[i] * C[(size_t)mu * nbf + i] * C[(size_t)nu * nbf + i];
            W[(size_t)mu * nbf + nu] = 2.0 * v;
        }
    std::vector<double> g2(n3);
    int rc = qcg_eri_gradient(c, D, 1.0, g2.data());
    if (rc) return rc;
    rc = qcg_one_electron_gradient(c, D, W.data(), grad);
    if (rc) return rc;
    for (int i = 0; i < n3; ++i) grad[i] += g2[i];
    return QCG_OK;
}

}  // extern "C"
