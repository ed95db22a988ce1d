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

int build_becke_grid(const HostBasis& B, const GridConfig& cfg, std::vector<double>& xyz, std::vector<double>& w,
                     std::vector<int>& atom_of);

}  // namespace qcg
