// Host side of the analytic nuclear gradient: the adjoint ERI kernels and the one-electron terms on
// the device, the nuclear-repulsion term on the host (rhf_gradient, scf/gradient.rs:900-976).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"

using namespace qcg;

namespace {

// Work items of the gradient kernels: every bra pair's surviving kets (the exact screening
// predicate Q_P Q_Q >= threshold, one segment per contraction bucket, as for J/K) cut into items
// of grad_item_kets(class pair) kets.  Built once per basis, threshold and primitive split and kept
// on the device.
int grad_plan(qcg_context* c) {
    const double key[2] = {c->schwarz_thr, (double)c->prim_split};
    if (c->d_grad_items && std::memcmp(key, c->grad_key, sizeof(key)) == 0) return QCG_OK;
    std::vector<int4> items;
    std::vector<int2> segs;
    int cls = 0;
    for (int cb = kNumPairClasses - 1; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++cls) {
            c->grad_off[cls] = (int)items.size();
            for (int p = 0; p < c->ket_cnt_h[cb]; ++p) {
                const int g0 = (int)segs.size();
                for_ket_ranges(c, cb, ck, 0, p, c->schwarz_thr, c->schwarz_thr,
                               [&](int bs, int, int e) { segs.push_back(make_int2(bs, e - bs)); });
                cut_items_split(items, segs, g0, p, grad_item_kets(cb, ck), c->pairs[c->class_off[cb] + p].npp,
                                c->prim_split);
            }
            c->grad_cnt[cls] = (int)items.size() - c->grad_off[cls];
        }
    c->release(c->d_grad_items);
    c->release(c->d_grad_segs);
    c->d_grad_items = nullptr;
    c->d_grad_segs = nullptr;
    HIPCHK(c, c->alloc(&c->d_grad_items, items.size()));
    HIPCHK(c, c->alloc(&c->d_grad_segs, segs.size()));
    HIPCHK(c, hipMemcpy(c->d_grad_items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_grad_segs, segs.data(), segs.size() * sizeof(int2), hipMemcpyHostToDevice));
    std::memcpy(c->grad_key, key, sizeof(key));
    return QCG_OK;
}

}  // namespace

// Launches the ERI gradient at the device density dD on st, ADDING into dgrad (natom x 3); the
// class kernels run on the auxiliary streams, forked from and joined back to st.  Returns without
// synchronising: the caller may do host work before it waits on st.
int qcg::eri_grad_launch(qcg_context* c, const double* dD, double hf_frac, double* dgrad, hipStream_t st,
                         const double* dW) {
    int rc = grad_plan(c);
    if (rc) return rc;
    rc = prepare_jk(c);  // the auxiliary streams and events
    if (rc) return rc;
    if (dW) {
        rc = onee_setup(c);
        if (rc) return rc;
    }
    HIPCHK(c, stream_begin(c, st));
    HIPCHK(c, hipEventRecord(c->ev_fork, st));
    for (int i = 0; i < qcg_context::kStreams; ++i) HIPCHK(c, hipStreamWaitEvent(c->aux[i], c->ev_fork, 0));
    // heaviest classes first: the d-bra classes with many components, by component count
    int order[21], cls_cb[21], cls_ck[21];
    for (int cb = kNumPairClasses - 1, i = 0; cb >= 0; --cb)
        for (int ck = cb; ck >= 0; --ck, ++i) {
            cls_cb[i] = cb;
            cls_ck[i] = ck;
            order[i] = i;
        }
    std::stable_sort(order, order + 21, [&](int x, int y) {
        return (double)c->grad_cnt[x] * c->class_nq[x] * (c->class_nq[x] > 27 ? 4 : 1) >
               (double)c->grad_cnt[y] * c->class_nq[y] * (c->class_nq[y] > 27 ? 4 : 1);
    });
    for (int k = 0; k < 21; ++k) {
        const int cls = order[k], cb = cls_cb[cls], ck = cls_ck[cls];
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
        a.items = c->d_grad_items + c->grad_off[cls];
        a.segs = c->d_grad_segs;
        a.nitems = c->grad_cnt[cls];
        const hipError_t e = launch_grad(cb, ck, a, c->aux[k % qcg_context::kStreams]);
        if (e != hipSuccess) return fail(c, QCG_DEVICE, std::string("launch_grad: ") + hipGetErrorString(e));
    }
    // the one-electron terms last, on the stream whose round-robin share of the (sorted) ERI
    // classes is the lightest
    if (dW) {
        const hipError_t e = launch_onee_grad(c->onee_view, c->d_onee_work, c->onee_nwork, dD, dW, c->d_rys, dgrad,
                                              c->aux[qcg_context::kStreams - 1]);
        if (e != hipSuccess) return fail(c, QCG_DEVICE, std::string("launch_onee_grad: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < qcg_context::kStreams; ++i) {
        HIPCHK(c, hipEventRecord(c->ev_join[i], c->aux[i]));
        HIPCHK(c, hipStreamWaitEvent(st, c->ev_join[i], 0));
    }
    HIPCHK(c, stream_done(c, st));
    return QCG_OK;
}

// W = 2 sum_i eps_i C_mu,i C_nu,i  (build_energy_weighted_density, scf/gradient.rs:101); rows over
// the host threads, the upper triangle mirrored.
std::vector<double> qcg::energy_weighted_density(int nbf, int n_occ, const double* C, const double* eps) {
    std::vector<double> W((size_t)nbf * nbf, 0.0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int mu = 0; mu < nbf; ++mu) {
        std::vector<double> ce(n_occ);
        for (int i = 0; i < n_occ; ++i) ce[i] = eps[i] * C[(size_t)mu * nbf + i];
        for (int nu = mu; nu < nbf; ++nu) {
            const double* cn = C + (size_t)nu * nbf;
            double v = 0.0;
            for (int i = 0; i < n_occ; ++i) v += ce[i] * cn[i];
            W[(size_t)mu * nbf + nu] = 2.0 * v;
        }
    }
    for (int mu = 0; mu < nbf; ++mu)
        for (int nu = 0; nu < mu; ++nu) W[(size_t)mu * nbf + nu] = W[(size_t)nu * nbf + mu];
    return W;
}

// Starts the ERI gradient of the host density D on c->stream into c->d_grad (zeroed first); with
// W (energy-weighted density, host) also the one-electron terms.
int qcg::eri_grad_start(qcg_context* c, const double* D, double hf_frac, const double* W) {
    const int nbf = c->basis.nbf, n3 = 3 * c->basis.natom;
    const size_t n2 = (size_t)nbf * nbf;
    if (!c->d_grad) HIPCHK(c, c->alloc(&c->d_grad, (size_t)n3));
    if (W && !c->d_W) HIPCHK(c, c->alloc(&c->d_W, n2));
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (W) HIPCHK(c, hipMemcpyAsync(c->d_W, W, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_fill(c->d_grad, n3, 0.0, c->stream));
    return eri_grad_launch(c, c->d_D, hf_frac, c->d_grad, c->stream, W ? c->d_W : nullptr);
}

// Waits for the device gradient and copies it to out (natom x 3).
int qcg::grad_finish(qcg_context* c, double* out) {
    HIPCHK(c, hipMemcpyAsync(out, c->d_grad, 3 * (size_t)c->basis.natom * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QCG_OK;
}

extern "C" {

int qcg_eri_gradient(qcg_context* c, const double* D, double hf_frac, double* grad) {
    if (!c || !D || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "gradient: one molecule per context");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = eri_grad_start(c, D, hf_frac);
    if (rc) return rc;
    return grad_finish(c, grad);
}

int qcg_rhf_gradient(qcg_context* c, const double* D, const double* C, const double* eps, int n_occ, double* grad) {
    if (!c || !D || !C || !eps || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "gradient: one molecule per context");
    const int nbf = c->basis.nbf, n3 = 3 * c->basis.natom;
    if (n_occ < 0 || n_occ > nbf) return fail(c, QCG_INVALID_ARG, "n_occ");
    HIPCHK(c, hipSetDevice(c->device));
    const std::vector<double> W = energy_weighted_density(nbf, n_occ, C, eps);
    // ERI and one-electron terms on the device, nuclear repulsion on the host meanwhile
    int rc = eri_grad_start(c, D, 1.0, W.data());
    if (rc) return rc;
    std::vector<double> g2(n3, 0.0);
    nuclear_repulsion_gradient(c->basis, g2.data());
    rc = grad_finish(c, grad);
    if (rc) return rc;
    for (int i = 0; i < n3; ++i) grad[i] += g2[i];
    return QCG_OK;
}

}  // extern "C"
