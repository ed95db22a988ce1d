// Becke/Lebedev molecular integration grid on the host (setup, FP64).
//
// Same grid as qc-core's build_becke_grid (crates/qc-core/src/dft/grid.rs:103-198):
//   * Mura-Knowles log3 radial rule, r = -R ln(1 - x^3), x = (i + 1/2)/n, R = 7.0 for
//     Li/Be/Na/Mg/K/Ca, else 5.2 (mura_knowles_radial :250-290);
//   * angular: SG-1 pruning (sg1_prune :315-345: Lebedev {6, 38, 86, 194, 86} by region of
//     r / R_atom against the row's alpha table) or the full rule (302 Standard / 590 Fine);
//   * Becke partition with three iterations of f(x) = x(3 - x^2)/2 and no atomic-size adjustment
//     (becke_partition :400-468); single-atom systems skip the partition;
//   * points with |w| <= 1e-30 are dropped (:177, :192).
// Points are emitted atom by atom, radial shell by shell, in Lebedev table order.
#include "grid.h"

#include <cmath>

#include "lebedev_table.inc"

namespace qcg {

namespace {

const double kSg1Radii[19] = {0.0,    1.0000, 0.5882, 3.0769, 2.0513, 1.5385, 1.2308, 1.0256, 0.8791, 0.7692,
                              0.6838, 4.0909, 3.1579, 2.5714, 2.1687, 1.8750, 1.6514, 1.4754, 1.3333};
const double kSg1Alphas[3][4] = {{0.25, 0.50, 1.00, 4.50}, {0.1667, 0.50, 0.90, 3.50}, {0.10, 0.40, 0.80, 2.50}};

int leb_index(int n) {
    for (int i = 0; i < 6; ++i)
        if (qcg_leb_orders[i] == n) return i;
    return -1;
}

}  // namespace

// Host Becke partition (the device one, becke_partition_kernel, repeats these operations in the
// same order).
static int host_partition(int64_t np, const double* pts, const int* owner, int nat, const double* axyz,
                          const double* dist, double* part) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < (long)np; ++p) {
#pragma clang fp contract(off)
        std::vector<double> gd(nat), P(nat, 1.0);
        for (int i = 0; i < nat; ++i) {
            const double dx = pts[3 * p] - axyz[3 * i], dy = pts[3 * p + 1] - axyz[3 * i + 1],
                         dz = pts[3 * p + 2] - axyz[3 * i + 2];
            gd[i] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        for (int i = 0; i < nat; ++i)
            for (int j = 0; j < i; ++j) {
                const double rij = dist[(size_t)i * nat + j];
                if (rij < 1e-15) continue;
                double mu = (gd[i] - gd[j]) / rij;
                for (int it = 0; it < 3; ++it) mu = mu * (3.0 - mu * mu) * 0.5;
                P[i] *= 0.5 * (1.0 - mu);
                P[j] *= 0.5 * (1.0 + mu);
            }
        double ps = 0.0;
        for (int i = 0; i < nat; ++i) ps += P[i];
        part[p] = std::fabs(ps) > 1e-30 ? P[owner[p]] / ps : 1.0 / nat;
    }
    return QCG_OK;
}

int build_becke_grid(const HostBasis& B, const GridConfig& cfg, std::vector<double>& xyz, std::vector<double>& w,
                     std::vector<int>& atom_of, BeckePartitionFn part_fn, void* user) {
    xyz.clear();
    w.clear();
    atom_of.clear();
    const int nat = B.natom;
    const int max_ang = cfg.quality == 0 ? 302 : 590;
    std::vector<double> dist((size_t)nat * nat, 0.0);
    for (int i = 0; i < nat; ++i)
        for (int j = i + 1; j < nat; ++j) {
            const double dx = B.xyz[3 * i] - B.xyz[3 * j], dy = B.xyz[3 * i + 1] - B.xyz[3 * j + 1],
                         dz = B.xyz[3 * i + 2] - B.xyz[3 * j + 2];
            dist[(size_t)i * nat + j] = dist[(size_t)j * nat + i] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    // all atoms' raw points (Mura-Knowles x Lebedev), then one partition over all of them
    std::vector<double> pts, raw;
    std::vector<int> owner;
    for (int a = 0; a < nat; ++a) {
        const int z = B.Z[a];
        const double rscale = (z == 3 || z == 4 || z == 11 || z == 12 || z == 19 || z == 20) ? 7.0 : 5.2;
        const int n = cfg.n_radial;
        for (int i = 0; i < n; ++i) {
            const double x = (i + 0.5) / (double)n;
            const double x3 = x * x * x;
            const double omx3 = 1.0 - x3;
            const double r = -rscale * std::log(omx3);
            const double dr = rscale * 3.0 * x * x / (omx3 * (double)n);
            int nang = max_ang;
            if (cfg.pruning) {
                const double ratom = (z < 19 ? kSg1Radii[z] : 1.0) + 1e-200;
                const int row = z <= 2 ? 0 : (z <= 10 ? 1 : 2);
                const double ratio = r / ratom;
                int place = 0;
                for (int k = 0; k < 4; ++k)
                    if (ratio > kSg1Alphas[row][k]) ++place;
                const int leb[5] = {6, 38, 86, 194, 86};
                nang = leb[place];
            }
            const int li = leb_index(nang);
            if (li < 0) return QCG_INVALID_ARG;
            const double* g = qcg_leb_xyzw + 4 * qcg_leb_off[li];
            const double rf = r * r * dr;
            for (int k = 0; k < nang; ++k) {
                pts.push_back(B.xyz[3 * a] + r * g[4 * k]);
                pts.push_back(B.xyz[3 * a + 1] + r * g[4 * k + 1]);
                pts.push_back(B.xyz[3 * a + 2] + r * g[4 * k + 2]);
                raw.push_back(rf * g[4 * k + 3]);
                owner.push_back(a);
            }
        }
    }
    const int64_t np = (int64_t)raw.size();
    std::vector<double> part(np, 1.0);
    if (nat > 1) {
        const int rc = part_fn ? part_fn(user, np, pts.data(), owner.data(), nat, B.xyz.data(), dist.data(), part.data())
                               : host_partition(np, pts.data(), owner.data(), nat, B.xyz.data(), dist.data(), part.data());
        if (rc) return rc;
    }
    for (int64_t p = 0; p < np; ++p) {
        const double wf = raw[p] * part[p];
        if (std::fabs(wf) > 1e-30) {
            xyz.push_back(pts[3 * p]);
            xyz.push_back(pts[3 * p + 1]);
            xyz.push_back(pts[3 * p + 2]);
            w.push_back(wf);
            atom_of.push_back(owner[p]);
        }
    }
    return QCG_OK;
}

}  // namespace qcg
