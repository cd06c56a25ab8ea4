// split.h -- host planning of the split form of the register-table PT kernel:
// one system shared by G workgroups ("parts"), each owning a strip of unit
// cells along the first lattice dimension plus a halo deep enough for one
// window of sweeps.
//
// A window is the span between two consecutive Metropolis sweeps of the
// parallel-tempering loop (monte_carlo.jl:295-305): nor overrelaxation
// sweeps then one Metropolis sweep, each a pass over the ncol colour classes,
// so P = (nor + 1) * ncol colour phases, followed by the total energy.  At
// the end of every window the parts swap the spins of their strip edges and
// their partial sums; in between each part updates, besides its own sites,
// every halo site whose new value one of its own sites will still depend on
// before the window ends.  That set shrinks phase by phase, so the sites of a
// colour are ordered (owned first, then by how late in the window they are
// still needed) and each phase updates a prefix of its colour class.
#pragma once
#include <array>
#include <vector>

#include "lattice.h"

namespace csmc {

struct SplitPart {
    int n_loc = 0, n_own = 0;
    std::vector<int> loc_dev;    // [n_loc] device site of each local site
    std::vector<int> col_start;  // [ncol + 1] local range of each colour class
    std::vector<int> own_cnt;    // [ncol] owned sites of each colour (range prefix)
    std::vector<int> phase_cnt;  // [P] sites updated at each phase (colour prefix)
    std::vector<int> nbl;        // [K2][n_loc] local neighbour index, n_loc if not local
    std::vector<int> bnd;        // owned local sites that another part holds (local order)
    std::vector<int> halo;       // local sites owned by another part, ordered by
                                 // (owner, position in the owner's bnd list)
    std::vector<int> halo_src;   // [n_halo] owner part and that position:
                                 // owner * 3 * maxbnd + position
    std::vector<std::array<int, 4>> slots;  // {first, colour end, basis, colour}
    std::vector<int> colslot;    // [ncol + 1] first slot of each colour
    long work = 0;               // wave updates per window (cost estimate)
};

struct SplitPlan {
    int G = 1, nor = 0, P = 0, ncol = 0, K2 = 0;
    int maxloc = 0, maxslots = 0, maxbnd = 0, maxhalo = 0;
    long work_max = 0;   // largest part's wave updates per window
    long work_full = 0;  // the same for the unsplit system
    std::vector<int> colbasis;  // [ncol] basis of each colour class
    std::vector<SplitPart> parts;
};

// Plan G parts for one window of nor overrelaxation sweeps and a Metropolis
// sweep.  Fails (returns false) when the lattice is too short along its
// first dimension or a colour class mixes bases.
bool build_split(const HostLattice &L, int G, int nor, int threads, SplitPlan &out);

// Replays the window phase by phase with a validity bit per local site,
// starting at every phase of the window, and checks that each part's own
// sites are correct after every phase and that the sites the window-end
// energy reads are correct at the end.  Returns the number of violations.
long check_split(const HostLattice &L, const SplitPlan &plan);

}  // namespace csmc
