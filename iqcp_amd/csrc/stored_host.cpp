// Preset-system path: the Fock build and the RHF SCF on a stored compressed two-electron tensor,
// for callers that hold S, H and the 8-fold tensor (PresetSystem, scf/mod.rs:284; its JSON form
// PresetSystemJson, :464) instead of a basis.  The context needs no basis.
#include <string>
#include <vector>

#include "context.h"
#include "scf.h"
#include "stored_launch.h"

using namespace qcg;

namespace {

// The size checks of PresetSystem::validate (scf/mod.rs:367-395): basis size and ERI length.
int validate_sizes(qcg_context* c, int nbf, int64_t eri_len) {
    if (nbf <= 0)
        return fail(c, QCG_INVALID_ARG, "Invalid number of basis functions: " + std::to_string(nbf) + " (must be positive)");
    const int64_t np = (int64_t)nbf * (nbf + 1) / 2, want = np * (np + 1) / 2;
    if (eri_len != want)
        return fail(c, QCG_INVALID_ARG, "Invalid ERI array size: expected " + std::to_string(want) + ", got " +
                                            std::to_string(eri_len));
    return QCG_OK;
}

// PresetSystem::validate (scf/mod.rs:359-395): electron count first, then the sizes.
int validate_stored(qcg_context* c, int nbf, int nelec, int64_t eri_len) {
    if (nelec <= 0 || nelec % 2)
        return fail(c, QCG_INVALID_ARG, "Invalid number of electrons: " + std::to_string(nelec) +
                                            " (must be positive and even for closed-shell RHF)");
    return validate_sizes(c, nbf, eri_len);
}

struct StoredUser {
    qcg_context* c;
    int nbf;
    const double* H;  // host
    double *dH, *dD, *dF, *deri;
};

int stored_fock_cb(void* u, const double* D, double* F, double* e_elec) {
    StoredUser* s = static_cast<StoredUser*>(u);
    qcg_context* c = s->c;
    const size_t n2 = (size_t)s->nbf * s->nbf;
    HIPCHK(c, hipMemcpyAsync(s->dD, D, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_fock_stored(s->nbf, s->dH, s->dD, s->deri, s->dF, c->stream));
    HIPCHK(c, hipMemcpyAsync(F, s->dF, n2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double e = 0.0;  // compute_electronic_energy (scf/mod.rs:898): 1/2 sum D (H + F)
    for (size_t i = 0; i < n2; ++i) e += D[i] * (s->H[i] + F[i]);
    *e_elec = 0.5 * e;
    return QCG_OK;
}

// Device copies of H and the tensor plus D, F work buffers; released by the caller.
int stored_buffers(qcg_context* c, int nbf, const double* H, const double* eri, int64_t eri_len, StoredUser* s) {
    const size_t n2 = (size_t)nbf * nbf;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->alloc(&s->dH, n2));
    HIPCHK(c, c->alloc(&s->dD, n2));
    HIPCHK(c, c->alloc(&s->dF, n2));
    HIPCHK(c, c->alloc(&s->deri, (size_t)eri_len));
    HIPCHK(c, hipMemcpyAsync(s->dH, H, n2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->deri, eri, (size_t)eri_len * sizeof(double), hipMemcpyHostToDevice, c->stream));
    return QCG_OK;
}

void release_buffers(qcg_context* c, StoredUser* s) {
    for (double* p : {s->dH, s->dD, s->dF, s->deri}) c->release(p);
}

}  // namespace

extern "C" {

int qcg_build_fock_stored(qcg_context* c, int nbf, const double* H, const double* D, const double* eri,
                          int64_t eri_len, double* F) {
    if (!c || !H || !D || !eri || !F) return QCG_INVALID_ARG;
    int rc = validate_sizes(c, nbf, eri_len);
    if (rc) return rc;
    StoredUser s{c, nbf, H, nullptr, nullptr, nullptr, nullptr};
    rc = stored_buffers(c, nbf, H, eri, eri_len, &s);
    if (rc == QCG_OK) {
        double e = 0.0;
        rc = stored_fock_cb(&s, D, F, &e);
    }
    release_buffers(c, &s);
    return rc;
}

int qcg_rhf_scf_stored(qcg_context* c, int nbf, int nelec, double e_nuc, const double* S, const double* H,
                       const double* eri, int64_t eri_len, const qcg_scf_config* cfg, const double* D_init,
                       qcg_scf_result* res, double* D, double* C, double* eps, double* F) {
    if (!c || !S || !H || !eri || !cfg) return QCG_INVALID_ARG;
    int rc = validate_stored(c, nbf, nelec, eri_len);
    if (rc) return rc;
    StoredUser s{c, nbf, H, nullptr, nullptr, nullptr, nullptr};
    rc = stored_buffers(c, nbf, H, eri, eri_len, &s);
    qcg_scf_result r{};
    if (rc == QCG_OK)
        rc = scf_driver(nbf, nelec / 2, S, H, e_nuc, *cfg, stored_fock_cb, &s, D_init, &r, D, C, eps, F, nullptr);
    release_buffers(c, &s);
    if (res) *res = r;
    if (rc == QCG_NOT_CONVERGED) return fail(c, rc, "ScfError::NotConverged");
    if (rc == QCG_NUMERICAL) return fail(c, rc, "ScfError::SingularOverlap");
    return rc;
}

}  // extern "C"
