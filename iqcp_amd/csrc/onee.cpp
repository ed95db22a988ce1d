// One-electron integrals on the device (setup work, FP64): overlap S, kinetic T, nuclear
// attraction V, and their nuclear derivatives for the gradient, with the per-element and
// per-primitive-pair bodies of onee_pair.h.
//
// Same matrices as qc-core's overlap_matrix (integrals/overlap.rs:647), kinetic (integrals/kinetic.rs),
// nuclear attraction (integrals/nuclear.rs:226, Rys quadrature) and hcore_matrix (integrals/hcore.rs:103),
// with the reference's per-primitive, per-Cartesian-component normalisation
// (cartesian_gaussian_normalization, integrals/overlap.rs:343-377) and raw contraction coefficients.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"
#include "onee_launch.h"

namespace qcg {

// nuclear_repulsion_gradient (scf/gradient.rs:135-176): dE/dR_A = -sum_B Z_A Z_B (R_A - R_B)/|R_A - R_B|^3
void nuclear_repulsion_gradient(const HostBasis& B, double* grad) {
    for (int a = 0; a < B.natom; ++a)
        for (int b = 0; b < B.natom; ++b) {
            if (a == b) continue;
            double d[3], r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                d[x] = B.xyz[3 * a + x] - B.xyz[3 * b + x];
                r2 += d[x] * d[x];
            }
            const double r = std::sqrt(r2);
            if (r < 1e-10) continue;
            for (int x = 0; x < 3; ++x) grad[3 * a + x] -= (double)B.Z[a] * B.Z[b] * d[x] / (r2 * r);
        }
}

namespace {

// Device copy of a basis: ints l, nprim, prim_off, shell_atom, bf_off (nshell each), Z (natom);
// doubles exps, coefs (one per primitive each), centers (3 per shell), xyz (3 per atom).
int upload_basis(qcg_context* c, const HostBasis& B, int** di, double** dd, BasisView* view) {
    std::vector<int> iv;
    for (const std::vector<int>* v : {&B.l, &B.nprim, &B.prim_off, &B.shell_atom, &B.bf_off, &B.Z})
        iv.insert(iv.end(), v->begin(), v->end());
    std::vector<double> dv;
    for (const std::vector<double>* v : {&B.exps, &B.coefs, &B.centers, &B.xyz}) dv.insert(dv.end(), v->begin(), v->end());
    HIPCHK(c, c->alloc(di, iv.size()));
    HIPCHK(c, c->alloc(dd, dv.size()));
    HIPCHK(c, hipMemcpy(*di, iv.data(), iv.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(*dd, dv.data(), dv.size() * sizeof(double), hipMemcpyHostToDevice));
    const int ns = B.nshell, np = (int)B.exps.size();
    const int* ip = *di;
    const double* dp = *dd;
    *view = BasisView{ns,      B.nbf,       B.natom,     ip,          ip + ns,         ip + 2 * ns, ip + 3 * ns,
                      ip + 4 * ns, ip + 5 * ns, dp,     dp + np,     dp + 2 * np,     dp + 2 * np + 3 * ns};
    return QCG_OK;
}

}  // namespace

// Device basis and (ordered shell pair, primitive pair) work list of the gradient kernel, made once
// per basis.
int onee_setup(qcg_context* c) {
    if (c->d_onee_work) return QCG_OK;
    const HostBasis& B = c->basis;
    int rc = upload_basis(c, B, &c->d_onee_int, &c->d_onee_dbl, &c->onee_view);
    if (rc) return rc;
    std::vector<int4> work;
    for (int si = 0; si < B.nshell; ++si)
        for (int sj = 0; sj < B.nshell; ++sj)
            for (int pi = 0; pi < B.nprim[si]; ++pi)
                for (int pj = 0; pj < B.nprim[sj]; ++pj) work.push_back(make_int4(si, sj, pi, pj));
    HIPCHK(c, c->alloc(&c->d_onee_work, work.size()));
    HIPCHK(c, hipMemcpy(c->d_onee_work, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice));
    c->onee_nwork = (int)work.size();
    return QCG_OK;
}

// S, T, V of the bases B[0..n) (any output may be null), full nbf x nbf row-major, on c's device:
// one allocation per kind for all of them, one launch per basis, one synchronisation.
int one_electron_device_many(qcg_context* c, int n, const HostBasis* const* B, double* const* S, double* const* T,
                             double* const* V) {
    std::vector<int> iv;
    std::vector<double> dv;
    std::vector<int4> work;
    std::vector<size_t> ioff(n), doff(n), woff(n + 1), moff(n + 1);
    moff[0] = 0;
    for (int m = 0; m < n; ++m) {
        const HostBasis& b = *B[m];
        ioff[m] = iv.size();
        doff[m] = dv.size();
        woff[m] = work.size();
        for (const std::vector<int>* v : {&b.l, &b.nprim, &b.prim_off, &b.shell_atom, &b.bf_off, &b.Z})
            iv.insert(iv.end(), v->begin(), v->end());
        for (const std::vector<double>* v : {&b.exps, &b.coefs, &b.centers, &b.xyz})
            dv.insert(dv.end(), v->begin(), v->end());
        for (int si = 0; si < b.nshell; ++si)
            for (int sj = si; sj < b.nshell; ++sj)
                for (int ia = 0; ia < ncart(b.l[si]); ++ia)
                    for (int jb = 0; jb < ncart(b.l[sj]); ++jb) work.push_back(make_int4(si, sj, ia, jb));
        moff[m + 1] = moff[m] + 3 * (size_t)b.nbf * b.nbf;
    }
    woff[n] = work.size();
    int* di = nullptr;
    double *dd = nullptr, *dm = nullptr;
    int4* dw = nullptr;
    int rc = QCG_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == QCG_OK) rc = fail(c, QCG_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(c->alloc(&di, iv.size()), "hipMalloc");
    chk(c->alloc(&dd, dv.size()), "hipMalloc");
    chk(c->alloc(&dw, work.size()), "hipMalloc");
    chk(c->alloc(&dm, moff[n]), "hipMalloc");
    if (rc == QCG_OK) {
        chk(hipMemcpyAsync(di, iv.data(), iv.size() * sizeof(int), hipMemcpyHostToDevice, c->stream), "upload");
        chk(hipMemcpyAsync(dd, dv.data(), dv.size() * sizeof(double), hipMemcpyHostToDevice, c->stream), "upload");
        chk(hipMemcpyAsync(dw, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream), "upload");
    }
    for (int m = 0; m < n && rc == QCG_OK; ++m) {
        const HostBasis& b = *B[m];
        const int ns = b.nshell, np = (int)b.exps.size();
        const int* ip = di + ioff[m];
        const double* dp = dd + doff[m];
        const BasisView view{ns,          b.nbf,       b.natom,     ip,      ip + ns,     ip + 2 * ns,
                             ip + 3 * ns, ip + 4 * ns, ip + 5 * ns, dp,      dp + np,     dp + 2 * np,
                             dp + 2 * np + 3 * ns};
        const size_t n2 = (size_t)b.nbf * b.nbf;
        double* mm = dm + moff[m];
        chk(launch_onee_matrices(view, dw + woff[m], (int)(woff[m + 1] - woff[m]), c->d_rys, mm, mm + n2, mm + 2 * n2,
                                 c->stream),
            "launch_onee_matrices");
        double* outs[3] = {S ? S[m] : nullptr, T ? T[m] : nullptr, V ? V[m] : nullptr};
        for (int k = 0; k < 3; ++k)
            if (outs[k])
                chk(hipMemcpyAsync(outs[k], mm + k * n2, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                    "download");
    }
    chk(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->release(di);
    c->release(dd);
    c->release(dw);
    c->release(dm);
    return rc;
}

int one_electron_device(qcg_context* c, const HostBasis& B, double* S, double* T, double* V) {
    const HostBasis* b = &B;
    return one_electron_device_many(c, 1, &b, S ? &S : nullptr, T ? &T : nullptr, V ? &V : nullptr);
}

}  // namespace qcg

using namespace qcg;

extern "C" {

int qcg_one_electron(qcg_context* c, double* S, double* T, double* V) {
    if (!c) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "one-electron matrices: one molecule per context");
    HIPCHK(c, hipSetDevice(c->device));
    return one_electron_device(c, c->basis, S, T, V);
}

int qcg_one_electron_gradient(qcg_context* c, const double* D, const double* W, double* grad) {
    if (!c || !D || !W || !grad) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    if (c->nmol != 1) return fail(c, QCG_INVALID_ARG, "gradient: one molecule per context");
    HIPCHK(c, hipSetDevice(c->device));
    const int nbf = c->basis.nbf, n3 = 3 * c->basis.natom;
    const size_t n2 = (size_t)nbf * nbf;
    int rc = onee_setup(c);
    if (rc) return rc;
    if (!c->d_grad) HIPCHK(c, c->alloc(&c->d_grad, (size_t)n3));
    if (!c->d_W) HIPCHK(c, c->alloc(&c->d_W, n2));
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_W, W, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_fill(c->d_grad, n3, 0.0, c->stream));
    HIPCHK(c, launch_onee_grad(c->onee_view, c->d_onee_work, c->onee_nwork, c->d_D, c->d_W, c->d_rys, c->d_grad,
                               c->stream));
    std::vector<double> g(n3, 0.0);
    nuclear_repulsion_gradient(c->basis, g.data());
    HIPCHK(c, hipMemcpyAsync(grad, c->d_grad, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n3; ++i) grad[i] += g[i];
    return QCG_OK;
}

}  // extern "C"
