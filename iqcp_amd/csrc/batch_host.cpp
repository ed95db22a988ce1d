// Many molecules iterating SCF together: one batched integral-direct J/K build per iteration for
// every molecule of the batch (quartet-class queues of all molecules in the same launches,
// captured once as a HIP graph and replayed each iteration), FP64 DIIS/diagonalisation per molecule
// on the host (the role of calling rhf_scf, scf/mod.rs:1363, once per molecule).
#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"
#include "scf.h"

using namespace qcg;

extern "C" {

int qcg_batch_build_jk(qcg_context* c, const double* D, double* J, double* K) {
    if (!c || !D) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n = (size_t)c->mat_total;
    HIPCHK(c, hipMemcpyAsync(c->d_D, D, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int rc = run_jk(c, c->d_D, J ? c->d_J : nullptr, K ? c->d_K : nullptr, c->stream);
    if (rc) return rc;
    if (J) HIPCHK(c, hipMemcpyAsync(J, c->d_J, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (K) HIPCHK(c, hipMemcpyAsync(K, c->d_K, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QCG_OK;
}

int qcg_batch_rhf_scf(qcg_context* c, const qcg_scf_config* cfg, int use_graph, qcg_scf_result* res, double* D_out,
                      double* ms_jk) {
    if (!c || !cfg || !res) return QCG_INVALID_ARG;
    if (!c->has_basis) return fail(c, QCG_NO_BASIS, "no basis");
    HIPCHK(c, hipSetDevice(c->device));
    const int nmol = c->nmol;
    std::vector<ScfRun> runs(nmol);
    std::vector<std::vector<double>> Ss(nmol), Hs(nmol);
    // one-electron matrices of all molecules on the device in one go
    {
        std::vector<std::vector<double>> Ts(nmol), Vs(nmol);
        std::vector<const HostBasis*> bp(nmol);
        std::vector<double*> sp(nmol), tp(nmol), vp(nmol);
        for (int m = 0; m < nmol; ++m) {
            const size_t n2 = (size_t)c->mols[m].nbf * c->mols[m].nbf;
            Ss[m].resize(n2);
            Hs[m].resize(n2);
            Ts[m].resize(n2);
            Vs[m].resize(n2);
            bp[m] = &c->mols[m];
            sp[m] = Ss[m].data();
            tp[m] = Ts[m].data();
            vp[m] = Vs[m].data();
        }
        const int r1 = one_electron_device_many(c, nmol, bp.data(), sp.data(), tp.data(), vp.data());
        if (r1) return r1;
        for (int m = 0; m < nmol; ++m)
            for (size_t i = 0; i < Hs[m].size(); ++i) Hs[m][i] = Ts[m][i] + Vs[m][i];
    }
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int m = 0; m < nmol; ++m) {
        const HostBasis& B = c->mols[m];
        int nelec = 0;
        for (int z : B.Z) nelec += z;
        if (nelec % 2) { bad |= 1; continue; }
        bad |= runs[m].init(B.nbf, nelec / 2, Ss[m].data(), Hs[m].data(), nuclear_repulsion(B), *cfg, nullptr) ? 1 : 0;
    }
    if (bad) return fail(c, QCG_INVALID_ARG, "batch: odd electron count or singular overlap");
    const size_t ntot = (size_t)c->mat_total;
    std::vector<double> Dp(ntot), Jp(ntot), Kp(ntot);
    // graph, exec and events are released on every return path below (GraphHold), after the
    // context stream has drained
    GraphHold gh;
    gh.st = c->stream;
    if (use_graph) {
        int rc = prepare_jk(c);  // plan and streams exist before capture: run_jk then only enqueues
        if (rc) return rc;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = run_jk(c, c->d_D, c->d_J, c->d_K, c->stream);
        hipError_t e = hipStreamEndCapture(c->stream, &gh.graph);
        if (rc) {
            (void)hipGetLastError();  // do not leave the capture error for the next call to find
            return rc;
        }
        HIPCHK(c, e);
        HIPCHK(c, hipGraphInstantiate(&gh.exec, gh.graph, nullptr, nullptr, 0));
    }
    HIPCHK(c, hipEventCreate(&gh.ev[0]));
    HIPCHK(c, hipEventCreate(&gh.ev[1]));
    double jk_ms = 0.0;
    int rc = QCG_OK;
    for (;;) {
        for (int m = 0; m < nmol; ++m)
            std::memcpy(&Dp[c->mat_off_h[m]], runs[m].D.data(), runs[m].D.size() * sizeof(double));
        HIPCHK(c, hipMemcpyAsync(c->d_D, Dp.data(), ntot * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipEventRecord(gh.ev[0], c->stream));
        if (use_graph) {
            HIPCHK(c, stream_begin(c, c->stream));
            HIPCHK(c, hipGraphLaunch(gh.exec, c->stream));
            HIPCHK(c, stream_done(c, c->stream));
        } else {
            rc = run_jk(c, c->d_D, c->d_J, c->d_K, c->stream);
            if (rc) break;
        }
        HIPCHK(c, hipEventRecord(gh.ev[1], c->stream));
        HIPCHK(c, hipMemcpyAsync(Jp.data(), c->d_J, ntot * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(Kp.data(), c->d_K, ntot * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, gh.ev[0], gh.ev[1]));
        jk_ms += ms;
        int active = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : active)
        for (int m = 0; m < nmol; ++m) {
            ScfRun& r = runs[m];
            if (r.done) continue;
            const int n = r.nbf;
            const size_t n2 = (size_t)n * n, o = (size_t)c->mat_off_h[m];
            std::vector<double> F(n2);
            double e = 0.0;
            for (size_t i = 0; i < n2; ++i) {
                F[i] = Hs[m][i] + Jp[o + i] - 0.5 * Kp[o + i];
                e += r.D[i] * (Hs[m][i] + F[i]);
            }
            if (!r.accept_fock(F.data(), 0.5 * e)) {
                r.step();
                ++active;
            }
        }
        if (active == 0) break;
    }
    if (rc) return rc;
    bool all = true;
    for (int m = 0; m < nmol; ++m) {
        runs[m].fill_result(&res[m]);
        all = all && runs[m].converged;
        if (D_out) std::memcpy(D_out + c->mat_off_h[m], runs[m].D.data(), runs[m].D.size() * sizeof(double));
    }
    if (ms_jk) *ms_jk = jk_ms;
    return all ? QCG_OK : fail(c, QCG_NOT_CONVERGED, "ScfError::NotConverged (batch)");
}

}  // extern "C"
