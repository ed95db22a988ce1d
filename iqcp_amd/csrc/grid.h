// Becke/Lebedev molecular grid (host).
#pragma once
#include <vector>

#include "onee.h"

namespace qcg {

struct GridConfig {
    int n_radial = 75;   // GridConfig::n_radial (dft/mod.rs:142-150)
    int quality = 0;     // 0 = Standard (Lebedev-302 region 4 / unpruned), 1 = Fine (590)
    int pruning = 1;     // SG-1 pruning
};

// Becke partition of np points: part[p] = P_owner[p](r_p) / sum_i P_i(r_p) (becke_partition,
// dft/grid.rs:400-468); dist = nat x nat internuclear distances.  Returns QCG_OK or an error code.
typedef int (*BeckePartitionFn)(void* user, int64_t np, const double* pts, const int* owner, int nat,
                                const double* atom_xyz, const double* dist, double* part);

// part_fn = nullptr: the partition runs on the host (OpenMP).
int build_becke_grid(const HostBasis& B, const GridConfig& cfg, std::vector<double>& xyz, std::vector<double>& w,
                     std::vector<int>& atom_of, BeckePartitionFn part_fn = nullptr, void* user = nullptr);

}  // namespace qcg
