// csmc_internal.h -- structures shared by the host runtime and the gfx950
// kernels of libcsmc.
//
// Device data layout (one HBM pool per kind, all lattices concatenated):
//   i32 pool   per lattice: nbr2[K2P][N], nbr3[K3][2][N], nbr4[K4][3][N]
//              (device site indices; an absent neighbour -- open boundary,
//              or one of the K2P - K2 padding slots -- is the pad site N,
//              whose spin is kept zero, so sweeps need no branch for it),
//              ref[N] (reference index of
//              each device site), sub[N] (basis of each device site, 0-based),
//              colrun[ncol+1] (run index boundaries per colour), incidence CSR
//              for the spin gradient, the sparsity pattern of each cubic and
//              quartic coupling.
//   run pool   int4 {start, end, sub, colour}: maximal ranges of device sites
//              that share colour and basis.  Device sites are ordered by
//              (colour, basis, reference index), so every colour class is one
//              contiguous range and each run has one coefficient set.
//              After the runs: the register-table slots, int4 {first site,
//              run end, basis, colour}, each run cut into REG_THREADS-site
//              pieces (thread t of a slot owns site first + t if < end).
//   coef pool  per lattice (float and double copies at the same offset):
//              onsite[nb][9], field[nb][3], c2[nb][K2P][9], c3[nb][K3][27],
//              c4[nb][K4][81]: the couplings of each (basis, slot) with the
//              role permutation of lattice.jl:184-193, 215-227, 251-272
//              already applied; padding slots are zero.
//   spins      packed {x,y,z} records of R per device site (12 B FP32,
//              24 B FP64) plus one zero pad site (N + 1 records per
//              system), systems back to back.
#pragma once
#include <stdint.h>

#include "../../include/csmc.h"

namespace csmc {

struct LatDev {
    int32_t N, nb, K2, K3, K4, ncol, nruns, has_self;
    int32_t K2P;         // bilinear slots padded to a multiple of 4
    int32_t flags;       // LAT_ONSITE | LAT_FIELD: some coupling is non-zero
    int64_t nbr_off;     // i32 pool
    int64_t ref_off;     // i32 pool
    int64_t sub_off;     // i32 pool
    int64_t colrun_off;  // i32 pool, run indices relative to run_off
    int64_t run_off;     // run pool (absolute index of first run)
    int64_t coef_off;    // coef pools
    int64_t inc_off;     // i32 pool: incidence offsets [N+1] then entries
    int64_t inc_n;       // number of incidence entries
    int64_t grad_off;    // i32 pool: slot->gradient index tables
    int32_t grad_len;    // coupling-gradient length per system
    int32_t nslots;      // register-table schedule: thread slots per sweep
    int64_t slot_off;    // run pool: int4 {first site, run end, basis, colour}
    int64_t colslot_off; // i32 pool: first slot of each colour [ncol + 1], then
                         // the basis all slots of a colour share (-1 if mixed) [ncol]
    int64_t pat_off;     // i32 pool: sparsity pattern (TP_*) of each cubic
                         // coupling [nb][K3], then of each quartic one [nb][K4]
    double S;
};

// LDS scratch at the start of every persistent sweep kernel's dynamic LDS:
// the workgroup reductions (block_sum of up to 9 values: 17 x 9 doubles,
// doubles 0..152), the split kernel's partial sums (RED_PART..RED_TOT + 8),
// then a few flags.
constexpr int RED_BYTES = 4096;
constexpr int RED_XCHG = 256;   // double index of the chain-exchange scratch (2)
constexpr int RED_FLAG = 258;   // double index of the exchange flag

// Workgroup size of the register-table sweep kernel; the slot schedule
// (one slot = REG_THREADS consecutive device sites of one run) is built for it.
constexpr int REG_THREADS = 1024;

// Split form of the register-table PT kernel: at most SPLIT_MAX workgroups
// per system.  LDS doubles RED_PART.. hold the parts' window partial sums
// (SPLIT_MAX x SPLIT_NV), RED_OWN.. this part's before they are published,
// RED_TOT.. their sum.
//
// Split tables (i32, one block per part, offsets from the table start):
// header [SPLIT_HDR] = n_loc, n_own, nslots, n_bnd, n_halo, then the offsets
// of col_start[ncol + 1], own_cnt[ncol], phase_cnt[P], slots (int4 {first,
// colour end, basis, colour}, 4-int aligned), colslot[ncol + 1] + colbasis
// [ncol], loc_dev[n_loc], loc_ref[n_loc], nbl[K2][n_loc], bnd[n_bnd] (owned
// sites another part holds), halo[n_halo][2] (local site, then where its
// owner publishes it: owner * 3 * maxbnd + position in the owner's bnd).
constexpr int SPLIT_MAX = 8;
constexpr int SPLIT_NV = 5;     // accepted count, energy, summed spin [3]
constexpr int RED_PART = 160;
constexpr int RED_OWN = 232;
constexpr int RED_TOT = 241;    // the summed window totals
constexpr int SPLIT_HDR = 16;   // header ints per part in the split tables

struct SysDev {
    int32_t lat;      // lattice id
    int32_t stream;   // Philox stream word of this system
    int64_t spin_off; // first spin record of this system (N + 1 records per system)
    int64_t site_off; // first site of this system in host-facing outputs
};

constexpr int32_t LAT_ONSITE = 1;
constexpr int32_t LAT_FIELD = 2;
constexpr int32_t LAT_DIAG = 4;  // every bilinear coupling matrix is diagonal
constexpr int32_t LAT_COLBASIS = 8;  // every colour class holds sites of one basis
// two colour classes, each of one basis and of the same whole number (1, 2
// or 4) of REG_THREADS-site register-table slots
constexpr int32_t LAT_TWO_COLOUR = 16;
// dE/dS is the local field (get_local_field, hamiltonian.jl:3-67): bilinear
// terms only, symmetric onsite matrices, no self-referencing sites, and every
// bond i -> j with matrix M matched by a bond j -> i with M^T
constexpr int32_t LAT_GRAD_LF = 32;
// some quartic slot names the site itself, and no bilinear or cubic slot
// does (a Metropolis step then evaluates only the quartic part twice)
constexpr int32_t LAT_SELF_Q4 = 64;
// some cubic or quartic slot is contracted in full (a TP_DENSE pattern)
constexpr int32_t LAT_DENSE_T = 128;

// Sparsity pattern of one (basis, slot) cubic or quartic coupling tensor,
// indices p (the site's own component), q, r, t (the partners in slot
// order).  A pattern names where the tensor may be non-zero; the field
// kernels then skip the products with the zero entries.  The terms that
// remain are added in the same order as in the full contraction, and a
// skipped product is an exact zero, so the sum is the same.
constexpr int32_t TP_DENSE = 0;
constexpr int32_t TP_ZERO = 1;    // every entry zero (a padding slot)
constexpr int32_t TP_C3_ALT = 2;  // cubic: p, q, r pairwise distinct (S.(S x S))
constexpr int32_t TP_Q4_PQ = 2;   // quartic: p = q and r = t ((Si.Sj)(Sk.Sl))
constexpr int32_t TP_Q4_PR = 3;   // quartic: p = r and q = t ((Si.Sk)(Sj.Sl))
constexpr int32_t TP_Q4_PT = 4;   // quartic: p = t and q = r ((Si.Sl)(Sj.Sk))

// Tiled sweep of a periodic Bravais lattice (tile.hip): the lattice is cut
// into tiles of T cells, each swept from a box of B cells (the tile plus the
// halo whose sites a whole sweep of the tile reads or recomputes) staged in
// LDS.  Box-local site index = b * boxcells + cell, cell last-dimension
// fastest.  Tables at tab_off in the tile table pool:
//   e[nb]                colour offset of each basis
//   runbase[p * nb]      device index of the first site of the run of raw
//                        colour r and basis b (-1: no such run)
//   delta[nb * K2]       box-local index step to the partner of slot k of
//                        basis b (TILE_PAD: the slot names no site)
//   rect[nsw * ncol * nb * 6]  box cells {lo, hi} per dimension where step
//                        s (sweep s / ncol, colour s % ncol) updates basis b
//                        (lo >= hi: none)
//   rawc[ncol]           raw colour (c . x + e_b) mod p of each colour
constexpr int32_t TILE_PAD = 0x7fffffff;

struct TileDev {
    int32_t D, nb, K2, p;
    int32_t cz;       // 1: the colour depends on the basis only
    int32_t cinv;     // inverse of c[D-1] mod p (cz == 0)
    int32_t ncol, ntiles;
    int32_t L[3], T[3], lo[3], B[3], nt[3], c[3];
    int32_t boxcells, nbox, tcells, ncell;
    int32_t tab_off;
    int32_t nsw;      // sweeps per launch (one staging of the box)
};

// incidence entry: owner device site and a packed (kind, slot, role)
//   code = kind(2 bits: 0 bil, 1 cub, 2 qua) | role << 2 (2 bits) | slot << 4
struct Inc {
    int32_t owner;
    int32_t code;
};

// Philox counter tags (fourth counter word)
constexpr uint32_t TAG_METRO = 0x4D455452u;  // "METR"
constexpr uint32_t TAG_EXCH = 0x45584348u;   // "EXCH"
constexpr uint32_t TAG_INIT = 0x494E4954u;   // "INIT"

// Parallel-tempering device state (one per PT batch).
struct PTDev {
    int32_t R;            // replicas (= temperature slots)
    int32_t alg;          // Metropolis kind
    int64_t t_th, t_meas, probe, swap, or_rate, dosweep;
    int32_t measure;
    int32_t pad_;
    uint64_t seed;
    // device arrays
    const double *temps;   // [R] temperature of slot
    int32_t *slot_of_sys;  // [R]
    int32_t *sys_of_slot;  // [R]
    double *E;             // [R] per system energy (FP64, last Metropolis sweep)
    double *sigma;         // [R] per slot
    double *acc_local;     // [R] per slot
    double *acc_exch;      // [R] per slot
    double *binners;       // [R][2][BIN_DOUBLES] per slot: E then |M|
    uint32_t *done;        // [1] WG completion counter for the exchange
    // chained segments (one cooperative launch, an exchange between each):
    // at exchange x the system in slot r publishes rec[x % PT_RECBUF][r] =
    // {E lo, E hi, sigma lo, sigma hi}, each word (x << 32 | payload), and
    // reads its partner slot's record; no other synchronisation
    uint64_t *rec;         // [PT_RECBUF][R][4]
    uint32_t *err;         // [1] non-zero: a wait ran out (no result)
    // probe series: (E, |M|) of each slot per probe sweep, folded into the
    // binners off the sweep path; [R][series_cap][2]
    double *series;
    int32_t series_cap;
    int32_t pad2_;
};


// records of the chained exchanges: a partner reads exchange x's record
// while the slot's next writers are at most a few exchanges ahead (each
// exchange waits on the partner slot, so neighbouring slots stay within two
// exchanges of each other); a record overwritten before it was read would
// carry another exchange index and its reader would time out, never read it
constexpr int PT_RECBUF = 8;

constexpr int BIN_LVLS = 32;
// per level: count, sum[2], sum2[4], pend[2], has  -> 10 doubles
constexpr int BIN_DOUBLES = BIN_LVLS * 10;

}  // namespace csmc
