// tile.h -- host planning of the tiled sweeps of periodic Bravais lattices
// (tile.hip): whether a lattice's sweeps can run from LDS tiles with
// neighbours found by index arithmetic, and the tile shape and halo.
#pragma once
#include <string>
#include <vector>

#include "csmc_internal.h"
#include "lattice.h"

namespace csmc {

struct TilePlan {
    bool ok = false;
    std::string why;             // why the lattice has no tiled form
    TileDev d{};                 // tab_off relative to tab
    std::vector<int32_t> tab;    // e, runbase, delta, rect, rawc
    double load_per_site = 0.0;  // box sites loaded per owned site
    double work_per_site = 0.0;  // updates (with halo recomputation) per owned site
};

// rec_bytes: LDS bytes per staged site; lds_budget: LDS bytes a tile may
// use; nsw: sweeps a launch runs from one staging of the box
TilePlan plan_tiles(const HostLattice &L, size_t rec_bytes, size_t lds_budget, int nsw = 1);

}  // namespace csmc
