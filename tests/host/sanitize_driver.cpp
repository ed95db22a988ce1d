// Host-only checks of the library's CPU code under AddressSanitizer / UndefinedBehaviorSanitizer:
// the SCF driver (scf.cpp, with DIIS, damping and level shift), the dense linear algebra
// (linalg.h) and the host Becke grid (grid.cpp).  No HIP call is made.  Built and run by
// tests/test_host_sanitizers.py; prints one "key value" line per result.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../iqcp_amd/csrc/grid.h"
#include "../../iqcp_amd/csrc/linalg.h"
#include "../../iqcp_amd/csrc/scf.h"

using namespace qcg;

double qcg::nuclear_repulsion(const HostBasis&) { return 0.0; }

namespace {

// A model closed-shell Fock operator F(D) = H + g * (D o M), M symmetric, positive entries:
// nonlinear in the occupied space, like a Coulomb term, so the SCF iterates.
struct Model {
    int n;
    std::vector<double> H, M;
    double g;
};

int model_fock(void* u, const double* D, double* F, double* e) {
    Model* m = static_cast<Model*>(u);
    const int n = m->n;
    double s = 0.0;
    for (int i = 0; i < n * n; ++i) {
        F[i] = m->H[i] + m->g * D[i] * m->M[i];
        s += D[i] * (m->H[i] + 0.5 * m->g * D[i] * m->M[i]);
    }
    *e = s;
    return QCG_OK;
}

}  // namespace

int main(int argc, char** argv) {
    const int natom_grid = argc > 1 ? std::atoi(argv[1]) : 6;
    std::mt19937_64 rng(7);
    std::normal_distribution<double> N(0.0, 1.0);
    // ---- linear algebra
    for (int n : {1, 2, 7, 40, 97}) {
        std::vector<double> A((size_t)n * n), w, V;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = N(rng);
        sym_eigen(A, n, w, V);
        double res = 0.0;
        for (int c = 0; c < n; ++c)
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * V[(size_t)k * n + c];
                res = std::max(res, std::fabs(s - w[c] * V[(size_t)i * n + c]));
            }
        std::printf("eigen_residual_%d %.3e\n", n, res);
        std::vector<double> b(n), x;
        for (double& v : b) v = N(rng);
        for (int i = 0; i < n; ++i) A[(size_t)i * n + i] += n;  // well conditioned
        const bool ok = lu_solve(A, b, n, x);
        double r2 = 0.0;
        for (int i = 0; i < n; ++i) {
            double s = -b[i];
            for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * x[k];
            r2 = std::max(r2, std::fabs(s));
        }
        std::printf("lu_ok_%d %d\nlu_residual_%d %.3e\n", n, ok ? 1 : 0, n, r2);
    }
    // ---- SCF driver on the model operator: DIIS, then damping + level shift
    const int n = 24, nocc = 5;
    Model m{n, std::vector<double>((size_t)n * n), std::vector<double>((size_t)n * n), 0.05};
    std::vector<double> S((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        S[(size_t)i * n + i] = 1.0;
        for (int j = 0; j < n; ++j) {
            const double h = (i == j) ? -2.0 + 0.15 * i : 0.2 * std::exp(-0.3 * std::abs(i - j));
            m.H[(size_t)i * n + j] = h;
            m.M[(size_t)i * n + j] = 1.0 / (1.0 + std::abs(i - j));
        }
        if (i + 1 < n) S[(size_t)i * n + i + 1] = S[(size_t)(i + 1) * n + i] = 0.1;
    }
    for (int variant = 0; variant < 2; ++variant) {
        ScfConfig cfg{1e-10, 1e-8, 200, 1, 6, 1, variant ? 0.3 : 0.0, 2, variant ? 0.2 : 0.0};
        ScfResult res{};
        std::vector<double> D((size_t)n * n), C((size_t)n * n), eps(n), F((size_t)n * n);
        const int rc = scf_driver(n, nocc, S.data(), m.H.data(), 0.0, cfg, model_fock, &m, nullptr, &res, D.data(),
                                  C.data(), eps.data(), F.data());
        // D S D = 2 D for a closed-shell density built from S-orthonormal orbitals
        double idem = 0.0, tr = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double s = 0.0;
                for (int k = 0; k < n; ++k)
                    for (int l = 0; l < n; ++l) s += D[(size_t)i * n + k] * S[(size_t)k * n + l] * D[(size_t)l * n + j];
                idem = std::max(idem, std::fabs(s - 2.0 * D[(size_t)i * n + j]));
                tr += D[(size_t)i * n + j] * S[(size_t)j * n + i];
            }
        std::printf("scf_rc_%d %d\nscf_converged_%d %d\nscf_energy_%d %.12f\nscf_idempotency_%d %.3e\nscf_electrons_%d %.12f\n",
                    variant, rc, variant, res.converged, variant, res.energy_total, variant, idem, variant, tr);
    }
    // ---- host Becke grid of a small water cluster (Z, xyz only)
    HostBasis B;
    const double pos[6][3] = {{0.0, 0.0, 0.0}, {1.43, 1.1, 0.0}, {-1.43, 1.1, 0.0},
                              {5.5, 0.3, 0.2}, {6.9, 1.4, 0.1}, {4.1, 1.4, -0.3}};
    for (int a = 0; a < natom_grid && a < 6; ++a) {
        B.Z.push_back(a % 3 == 0 ? 8 : 1);
        for (int x = 0; x < 3; ++x) B.xyz.push_back(pos[a][x]);
    }
    B.natom = (int)B.Z.size();
    for (int pruning = 0; pruning < 2; ++pruning) {
        GridConfig g;
        g.pruning = pruning;
        std::vector<double> xyz, w;
        std::vector<int> at;
        const int rc = build_becke_grid(B, g, xyz, w, at);
        double wsum = 0.0;
        for (double v : w) wsum += v;
        std::printf("grid_rc_%d %d\ngrid_points_%d %zu\ngrid_wsum_%d %.6f\n", pruning, rc, pruning, w.size(), pruning,
                    wsum);
    }
    return 0;
}
