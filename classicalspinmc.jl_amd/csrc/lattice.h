// lattice.h -- host-side Lattice construction for libcsmc: the per-site
// interaction tables of ClassicalSpinMC.jl's Lattice constructor
// (lattice.jl:65-291), compacted per basis, plus the hypergraph colouring
// and the colour-major device ordering.
#pragma once
#include <array>
#include <string>
#include <vector>

#include "../../include/csmc.h"

namespace csmc {

// A lattice-periodic colouring colour(b, x) = remap[(c . x + e_b) mod p]
// (p = 0: the colouring is not of this form)
struct ColourPattern {
    int p = 0;
    int c[3] = {0, 0, 0};
    std::vector<int> e, remap;
};

struct HostLattice {
    int dim = 0, nb = 0, periodic = 1;
    int shape[3] = {1, 1, 1};
    double S = 0.5;
    long N = 0, ncell = 0;

    // unit-cell copy (term order = call order)
    std::vector<double> avec, basis;
    std::vector<int> fb, ob, b2, b3, b4, off2, off3, off4;
    std::vector<double> fh, om, m2, t3, t4;
    int n2 = 0, n3 = 0, n4 = 0;

    // field / onsite source call per basis (-1 = zero), lattice.jl:120-140
    std::vector<int> fsrc, osrc;

    // compaction: for basis b the active terms in term order
    int K2 = 0, K3 = 0, K4 = 0;
    std::vector<int> slot2, slot3, slot4;  // [nb][K] term id or -1
    std::vector<int> role2, role3, role4;  // [nb][K] role of the site in the term
    // neighbours in reference order, 0-based, -1 = absent
    std::vector<int> nbr2;  // [N][K2]
    std::vector<int> nbr3;  // [N][K3][2]
    std::vector<int> nbr4;  // [N][K4][3]
    // couplings per basis and slot, role permutation applied
    std::vector<double> coef;
    long coef_onsite = 0, coef_field = 0, coef_c2 = 0, coef_c3 = 0, coef_c4 = 0, coef_len = 0;

    // colouring and device order
    int ncol = 0;
    std::vector<int> colour;                 // [N] reference order
    std::vector<int> dev_of_ref, ref_of_dev;  // permutations
    std::vector<std::array<int, 4>> runs;    // {start, end, basis, colour}
    std::vector<int> colrun;                 // [ncol+1] run boundaries
    ColourPattern pattern;                   // how `colour` was made, if periodic
    int has_self = 0;

    // gradient layout: field calls*3, onsite calls*9, n2*9, n3*27, n4*81
    int grad_len = 0;

    int basis_of(long site) const { return (int)(site / ncell); }
};

// Returns CSMC_OK or an error code with `err` set.
int build_lattice(const csmc_unit_cell *uc, const int32_t *shape, int periodic, double S,
                  HostLattice &L, std::string &err);
// Recompute couplings (coef) from a unit cell with identical term structure.
int update_couplings(const csmc_unit_cell *uc, HostLattice &L, std::string &err);
// Reference-layout tables (lattice.jl:12-22).
void export_tables(const HostLattice &L, double *pos, double *field, double *onsite,
                   int64_t *bsite, double *bmat, int64_t *csite, double *ctens, int64_t *qsite,
                   double *qtens);
// Colour the interaction hypergraph; fills colour/ncol and the device order.
void colour_lattice(HostLattice &L);
// true if no two sites of one colour share a term (host validity check)
bool colouring_valid(const HostLattice &L, std::string &err);

}  // namespace csmc
