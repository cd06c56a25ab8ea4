// kernels.h -- launch interface between the libcsmc host runtime and its
// gfx950 kernels.
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csmc_internal.h"

namespace csmc {

struct DevPools {
    const LatDev *lats;
    const int32_t *i32;
    const int4 *runs;
    const float *coefF;
    const double *coefD;
    const SysDev *sys;
};

// WG_PT runs one segment per launch (the last workgroup does the exchange);
// WG_PT_CHAIN runs a list of segments in one cooperative launch, each exchange
// decided pair by pair from records the partner slots publish.
enum WGMode { WG_SWEEPS = 0, WG_ANNEAL = 1, WG_PT = 2, WG_PT_CHAIN = 3 };

// Split form of the chained PT kernel: workgroup b runs part b / R of
// system b % R.  Parts swap strip edges and partial sums through words of
// (epoch << 32 | payload), each written and read as one 64-bit atomic, so a
// reader that sees the epoch it waits for also sees the payload.
struct SplitDev {
    int32_t G;            // parts per system (1: the whole-system kernel)
    int32_t nor;          // overrelaxation sweeps per window
    int32_t R;            // systems
    int32_t bmax;         // most published (boundary) sites of a part
    const int32_t *tab;   // split tables (csmc_internal.h)
    uint64_t *hx;         // [2][R][G][3][bmax] published spin words (x, y, z
                          // planes of each part), by epoch parity
    uint64_t *px;         // [2][R][SPLIT_MAX][2 * SPLIT_NV] partial-sum words
    uint32_t *epoch;      // [R] epoch of each system's next exchange
};

// Arguments of the persistent workgroup-per-system kernel.
struct WGArgs {
    DevPools P;
    void *spins;
    int mode;
    int kind;           // sweep kind (WG_SWEEPS) or Metropolis alg (anneal / PT)
    uint32_t k0, k1;    // Philox key (seed)
    uint32_t *mcount;   // [nsys] Metropolis sweep counters
    // WG_SWEEPS
    int n_sweeps;
    const double *T_sys;
    double *sigma_sys;
    double *acc_sys;
    // WG_ANNEAL
    const double *temps;
    int n_temps;
    int64_t t_th, or_rate;
    double sigma0;
    double *acc_T;      // [nsys][n_temps]
    // WG_PT
    PTDev pt;
    int64_t seg_begin, seg_end;
    int begin_mid, end_mid, exchange;
    // WG_PT_CHAIN: nseg segments, the first from seg_begin (begin_mid) to the
    // exchange at chain_e0, segment g to the exchange at chain_e0 + g * chain_P
    int64_t chain_e0, chain_P;
    int64_t chain_tail;  // >= 0: after the last exchange, run on to the end of
                         // sweep chain_tail without a further exchange
    int nseg;
    uint32_t spin_limit;  // polls of a wait for another workgroup before it gives up
    int64_t probe_base;   // series index of this launch's first probe
    SplitDev split;       // WG_PT_CHAIN with split.G > 1
    // WG_PT_CHAIN in storage mode 0: parts0 workgroups per system (block b
    // is part b % parts0 of system b / parts0); bar0 [nsys][2] holds each
    // system's barrier (arrivals, generation), psum0 [nsys][2][parts0][8]
    // the parts' workgroup sums
    int parts0;
    uint32_t *bar0;
    double *psum0;
    uint32_t *err0;  // set when a barrier wait of the parts gives up
    uint32_t *xcc0;  // [nsys][parts0]: the XCD each part runs on, written at launch
    int ptab;        // storage mode 0: stage each workgroup's partners in LDS (the
                     // launch's LDS bytes hold the table)
    // WG_SWEEPS / WG_ANNEAL in storage mode 0 with a packed schedule:
    // workgroup w runs tasks[task_off[w] .. task_off[w + 1] - 1] in turn,
    // each {system, part, parts of that system, 0}; psum0 and xcc0 are laid
    // out with parts0 parts per system.  Null: workgroup w is part w % parts0
    // of system w / parts0.
    const int32_t *task_off;
    const int4 *tasks;
};

// Arguments of the one-colour-pass kernel (plain / stream form).
struct PassArgs {
    DevPools P;
    void *spins;
    int sys0;
    int colour;
    int kind;
    uint32_t k0, k1;
    const uint32_t *ctr;     // [nsys] Philox sweep counter word per system, or null
    uint32_t ctr0;           // counter word used when ctr is null
    const double *T_sys;
    const double *sigma_sys;
    uint32_t *acc;           // [nsys] accepted proposals
};

// one whole coloured sweep of every system from LDS tiles (tile.hip)
struct TileArgs {
    DevPools P;
    const void *spins_in;    // the state before the sweep
    void *spins_out;         // the state after it (every site is written)
    const TileDev *td;       // [lattice id] tile plan (ntiles 0: not used)
    const int32_t *tab;      // tile table pool
    int flat;                // every plan's first position has extent 1
    int sys0;
    int kind;
    uint32_t k0, k1;
    const uint32_t *ctr;     // [nsys] Philox sweep counter word per system, or null
    uint32_t ctr0;           // counter word used when ctr is null
    const double *T_sys;
    const double *sigma_sys;
    uint32_t *acc;           // [nsys] accepted proposals
};

// metropolis_constraint_adaptive! sweeps (constraint.hip) for the constraint
// c = sum_i (a_b . s_i + s_i . Q_b s_i) - c0
struct ConsArgs {
    DevPools P;
    void *spins;
    int colour;
    uint32_t k0, k1;
    uint32_t *ctr;          // [nsys] Philox Metropolis sweep counters
    const double *T_sys;    // [nsys]
    double *sigma_sys;      // [nsys] cone widths (updated after a sweep)
    const double *coef;     // [cnb][12]: a_b (3) then Q_b (9, row-major)
    int cnb;
    double c0, weight, sigma0;
    void *prop;             // [spin slots] proposed spin records
    double *dE, *dC, *u;    // [spin slots]
    uint8_t *dec;           // [spin slots] decisions
    double *lambda, *cval;  // [nsys]
    double *acc_sweep;      // [nsys] accepted proposals of the current sweep
    double *acc_total;      // [nsys] accepted proposals since the call began
};

// c of every system and a cleared sweep count, at the start of each sweep
hipError_t launch_cons_begin(int prec, int nsys, const ConsArgs &a, hipStream_t st);
// one colour pass (a.colour) of every system: proposals, the per-system
// chain of decisions, the stores
hipError_t launch_cons_pass(int prec, bool self, bool gen, int nsys, int max_colour_sites, const ConsArgs &a,
                            hipStream_t st);

int max_lds_bytes();
// threads per workgroup of the persistent sweep kernel for a batch of this
// precision, with cubic / quartic / self terms (general), all in sparse
// patterns (sparse)
int wg_threads_for(int prec, bool general, bool sparse);
// raise a kernel's dynamic LDS limit to the device maximum (once per kernel
// and device; later calls return at once)
hipError_t allow_large_lds(const void *fn);
// hipLaunchCooperativeKernel of a persistent kernel taking WGArgs (fails when
// the grid cannot be resident all at once)
hipError_t launch_cooperative(const void *fn, int nblocks, int threads, size_t lds_bytes,
                              const WGArgs &a, hipStream_t st);
// lm: storage mode (0 global, 1 LDS spins, 2 LDS spins + uint16 neighbour
// tables); self / gen: some lattice of the batch has a self-referencing
// term / cubic or quartic terms; k2c: the bilinear slot count shared by
// every lattice of the batch (0 if they differ)
// sparse: no cubic or quartic slot of the batch's lattices is contracted in
// full (TP_DENSE), so the FP32 general kernels may leave the full
// contractions out
hipError_t launch_wg(int prec, int lm, bool self, bool gen, int k2c, bool sparse, int nsys, size_t lds_bytes,
                     const WGArgs &a, hipStream_t st);
// register-table mode (FP32, bilinear-only): k2c in {3, 4, 6}, spt in {4, 8}
// register slots per thread (8 only with k2c <= 4), diag: every coupling
// matrix of the batch is diagonal; two_colour: every lattice has two colour
// classes of spt / 2 full REG_THREADS-site slots each (LAT_TWO_COLOUR);
// lds_bytes = 16 * (max sites + 1)
hipError_t launch_wg_reg(int k2c, int spt, bool diag, bool two_colour, int nsys, size_t lds_bytes,
                         const WGArgs &a, hipStream_t st);
// split register-table chain (WG_PT_CHAIN only): G * R workgroups,
// lds_bytes = 16 * (most local sites of a part + 1) + 4 * (most boundary
// sites) + 8 * (most halo sites)
hipError_t launch_wg_split(int k2c, int spt, bool diag, int G, int R, size_t lds_bytes,
                           const WGArgs &a, hipStream_t st);
// grid (grid_x, nsys): workgroup x of system y sweeps tiles x, x + grid_x, ...
hipError_t launch_tile(int prec, bool self, int nsys, int grid_x, size_t lds_bytes, const TileArgs &a,
                       hipStream_t st);
hipError_t launch_pass(int prec, bool self, bool gen, int nsys, int max_colour_sites,
                       const PassArgs &a, hipStream_t st);
// captured plain-form sweeps: the per-sweep counter (and adaptive cone)
// step of systems s0 .. s0 + count - 1 (adaptive: count 1), and the clearing
// of the accepted counters
hipError_t launch_graph_after_sweep(uint32_t *ctr, uint32_t *acc, double *total, double *sigma, int s0,
                                    int count, bool adaptive, double n_sites, hipStream_t st);
hipError_t launch_graph_clear(uint32_t *acc, double *total, int nsys, hipStream_t st);
hipError_t launch_site_eval_at(int prec, const DevPools &P, void *spins, int sys, const int32_t *set_dev,
                               const double *set_val, int n_set, const int32_t *at_dev, int n_at, double *H,
                               double *e, hipStream_t st);
hipError_t launch_site_eval(int prec, const DevPools &P, const void *spins, int sys, int N,
                            double *H, double *e, hipStream_t st);
hipError_t launch_energy(int prec, const DevPools &P, const void *spins, int nsys, double *E,
                         double *parts, hipStream_t st);
hipError_t launch_magnet(int prec, const DevPools &P, const void *spins, int nsys, double *M,
                         hipStream_t st);
hipError_t launch_randomize(int prec, const DevPools &P, void *spins, int nsys, int maxN,
                            uint32_t k0, uint32_t k1, hipStream_t st);
// src: double[N][3] in reference order (device buffer)
hipError_t launch_scatter_spins(int prec, const DevPools &P, void *spins, int sys, int N,
                                const double *src, hipStream_t st);
hipError_t launch_gather_spins(int prec, const DevPools &P, const void *spins, int sys, int N,
                               double *dst, hipStream_t st);
// dEdS: double[total sites][3], systems concatenated in batch order
hipError_t launch_grad_spins(int prec, const DevPools &P, const void *spins, int nsys, int maxN,
                             double *dEdS, hipStream_t st);
hipError_t launch_grad_couplings(int prec, const DevPools &P, const void *spins, int nsys,
                                 int grad_len, double *dEdJ, hipStream_t st);
// push probes [0, n) of every slot's series into its binners, in order
hipError_t launch_bin_fold(int R, const double *series, int cap, int n, double *binners,
                           hipStream_t st);
// one fused gradient step from spins_old into spins_new (same layout)
hipError_t launch_gd_step(int prec, const DevPools &P, const void *spins_old, void *spins_new, int nsys, int maxN,
                          double eta, hipStream_t st);
hipError_t launch_gd(int prec, const DevPools &P, void *spins, int nsys, int maxN,
                     const double *dEdS, double eta, hipStream_t st);
// equal-time correlations of nsys systems of one lattice (N sites, device
// positions pos[N][3]) at momenta ks[nk][3], in blocks of kb momenta:
// workspaces C, Q [kb][N], B [3 nsys][N], Cre, Cim [3 nsys][kb];
// out[nsys][nk][9]
// equal-time correlations, fused FP64 MFMA path: sizes of the padded
// operands and of the split of the site sum
constexpr int SSF_KC = 32;  // sites per LDS chunk
// the fused path's own branch-free sincos is used when every |k.r| is
// below this bound (OCML's sincos otherwise)
constexpr double SSF_FAST_ARG = 1e5;
struct SsfPlan {
    int N, nk, nsys;
    int ph;                   // phases: 0 library sincos, 1 branch-free
                              // sincos, 2 rotations by exp(i k.d)
    double d[3];              // the step between sites 4 apart (ph 2)
    int nct;                  // 16-column tiles per wave (1, 2, 3, 4, 6, 8, 12)
    int colblocks, ldb;       // column blocks of 16 nct; ldb = colblocks 16 nct
    int mblocks, mpad;        // blocks of 64 momenta; mpad = 64 mblocks
    int nchunks, npad;        // chunks of SSF_KC sites; npad = SSF_KC nchunks
    int64_t units;            // (tile, chunk) pairs, tiles = mblocks colblocks
    int grid;                 // workgroups: one round of resident ones
    int maxparts;             // most workgroups sharing one tile
    std::vector<int32_t> wfirst, nparts;  // per tile: first workgroup, count
};
void ssf_plan(int N, int nk, int nsys, int ncu, int ph, SsfPlan *p);
// perm: N device site indices in the order of pos; tiles: wfirst then
// nparts (2 x mblocks colblocks); pos: npad x 3, ks: mpad x 3
// (zero-padded); B: npad x ldb and P: tiles x maxparts x 2 x 64 x 16 nct
// scratch; out: nsys x nk x 9; ev0 / ev1 (may be null) are recorded around
// the MFMA kernel alone
hipError_t launch_ssf_mfma(const SsfPlan &p, int prec, const void *spins, const SysDev *sys, const int32_t *perm,
                           const int32_t *tiles, const double *pos, const double *ks, double *B, double *P,
                           double *out, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);
// acc[T][k][j] += out[sys_of_slot[T]][k][j] for the R temperature slots
hipError_t launch_ssf_accumulate(const double *out, const int32_t *sys_of_slot, int R, int nk, double *acc,
                                 hipStream_t st);
// the rocBLAS path: phase matrices in memory, two dgemm per block of momenta
hipError_t launch_ssf_blas(void *blas, int prec, const void *spins, const SysDev *sys, int nsys, int N,
                      const double *pos, const double *ks, int nk, int kb, double *C, double *Q,
                      double *B, double *Cre, double *Cim, double *out, hipStream_t st);
hipError_t blas_create(void **blas);
void blas_destroy(void *blas);

}  // namespace csmc
