/*
 * csmc.h -- C ABI of libcsmc, the MI355X (gfx950) sweep engine for
 * ClassicalSpinMC.jl-style classical spin Monte Carlo.
 *
 * Conventions (chosen so a Julia ccall can pass the reference's own arrays):
 *   - every entry point returns int32 status, CSMC_OK == 0; the message of the
 *     last failure on the calling thread is csmc_last_error();
 *   - host pointers are borrowed for the duration of the call only;
 *   - basis indices are 1-based (unit_cell.jl:23-75), site indices exported in
 *     reference tables are 1-based with 0 = absent (lattice.jl:16-21);
 *   - spins are double[N][3] == Julia's 3xN column-major Array{Float64,2}
 *     (lattice.jl:6), in the reference's site order (lattice.jl:29-33);
 *   - InteractionMatrix is row-major m11..m33 (interaction_matrix.jl:2-12);
 *     cubic / quartic tensors are Julia column-major Array{Float64,3/4}:
 *     C[a,b,c] at a + 3b + 9c, R[a,b,c,d] at a + 3b + 9c + 27d (0-based).
 *   - calls on one context are serialised on that context's HIP stream and
 *     are not re-entrant; use one context per host thread.
 */
#ifndef CSMC_H
#define CSMC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSMC_OK 0
#define CSMC_EINVAL 1   /* invalid argument                                    */
#define CSMC_EBOUNDS 2  /* the reference raises BoundsError (lattice.jl:132,138) */
#define CSMC_EUNDEF 3   /* the reference reads an undef entry (lattice.jl:122,126) */
#define CSMC_EBC 4      /* error("Invalid boundary condition option") (lattice.jl:107) */
#define CSMC_EHIP 5     /* HIP runtime failure                                  */
#define CSMC_ENOMEM 6
#define CSMC_ESTATE 7   /* object in the wrong state / unknown handle           */

#define CSMC_FP32 32
#define CSMC_FP64 64

/* sweep kinds; the Metropolis kinds are the alg plugin of
 * metropolis.jl:181-199 */
#define CSMC_OVERRELAX 0       /* overrelaxation!        monte_carlo.jl:126-139 */
#define CSMC_METROPOLIS 1      /* metropolis!            metropolis.jl:65-82    */
#define CSMC_CONE 2            /* metropolis_fixed_cone! metropolis.jl:136-153  */
#define CSMC_CONE_ADAPTIVE 3   /* metropolis_adaptive!   metropolis.jl:113-133  */
#define CSMC_DETERMINISTIC 4   /* deterministic_updates! monte_carlo.jl:201-213 */
/* metropolis_constraint_adaptive! (metropolis.jl:155-179) for the constraint
 * set by csmc_constraint_set: csmc_sweep and csmc_anneal only */
#define CSMC_CONSTRAINT_ADAPTIVE 5

typedef struct csmc_ctx csmc_ctx;
typedef struct csmc_graph csmc_graph;

/* UnitCell{D} (unit_cell.jl:3-18) flattened.  Term arrays are in call order
 * of addZeemanCoupling!/addOnSite!/addBilinear!/addCubic!/addQuartic!. */
typedef struct csmc_unit_cell {
    int32_t dim;                    /* D in 1..3                               */
    int32_t n_basis;
    const double *lattice_vectors;  /* [D][D]: vector d at d*D                 */
    const double *basis;            /* [n_basis][D]                            */
    int32_t n_field;
    const int32_t *field_b;         /* [n_field]                               */
    const double *field_h;          /* [n_field][3]                            */
    int32_t n_onsite;
    const int32_t *onsite_b;        /* [n_onsite]                              */
    const double *onsite_m;         /* [n_onsite][9] row-major                 */
    int32_t n_bilinear;
    const int32_t *bilinear_b;      /* [n][2]  (b1, b2)                        */
    const double *bilinear_m;       /* [n][9]  row-major                       */
    const int32_t *bilinear_off;    /* [n][D]                                  */
    int32_t n_cubic;
    const int32_t *cubic_b;         /* [n][3]                                  */
    const double *cubic_t;          /* [n][27] column-major                    */
    const int32_t *cubic_off;       /* [n][2][D]  (o2, o3)                     */
    int32_t n_quartic;
    const int32_t *quartic_b;       /* [n][4]                                  */
    const double *quartic_t;        /* [n][81] column-major                    */
    const int32_t *quartic_off;     /* [n][3][D]  (o2, o3, o4)                 */
} csmc_unit_cell;

typedef struct csmc_lattice_info {
    int64_t n_sites;       /* Lattice.size (lattice.jl:10)                     */
    int32_t dim, n_basis;
    int32_t n2, n3, n4;    /* per-site tuple widths N2, N3, N4 (lattice.jl:4)   */
    int32_t k2, k3, k4;    /* device slots per site after compaction           */
    int32_t n_colours;     /* independent sets of the interaction hypergraph   */
    int32_t has_self;      /* some term of a site references the site itself   */
    double S;
} csmc_lattice_info;

/* Parallel-tempering schedule: SimulationParameters (metropolis.jl:4-13). */
typedef struct csmc_pt_params {
    int64_t t_thermalization, t_measurement, probe_rate, swap_rate;
    int64_t overrelaxation_rate, report_interval, checkpoint_rate;
    int32_t alg;            /* CSMC_METROPOLIS / CSMC_CONE / CSMC_CONE_ADAPTIVE */
    int32_t measure;        /* 1: accumulate E and |M| binning per slot         */
} csmc_pt_params;

/* ---- library / context -------------------------------------------------- */
int32_t csmc_version(void);
const char *csmc_last_error(void);
/* no reference counterpart: device discovery for the HIP backend */
int32_t csmc_device_count(int32_t *n);
int32_t csmc_ctx_create(int32_t device, csmc_ctx **out);
int32_t csmc_ctx_destroy(csmc_ctx *ctx);
int32_t csmc_ctx_sync(csmc_ctx *ctx);
/* no reference counterpart: kernel timing with HIP events recorded on the
 * context's stream around every sweep-kernel launch (enable != 0 starts a
 * fresh window); read returns the summed kernel time and the launch count. */
int32_t csmc_timing(csmc_ctx *ctx, int32_t enable);
int32_t csmc_timing_read(csmc_ctx *ctx, double *total_ms, int64_t *launches);

/* ---- Lattice ------------------------------------------------------------ */
/* Host-only form of the Lattice constructor (lattice.jl:65-291): builds the
 * tables and the colouring without a device.  info receives the sizes; any
 * output pointer may be NULL (layouts as csmc_lattice_tables; colour[N]). */
int32_t csmc_build_tables(const csmc_unit_cell *uc, const int32_t *shape, int32_t periodic,
                          double S, csmc_lattice_info *info, double *pos, double *field,
                          double *onsite, int64_t *bsite, double *bmat, int64_t *csite,
                          double *ctens, int64_t *qsite, double *qtens, int32_t *colour);
/* Plan of the split parallel-tempering kernel for a lattice: `parts`
 * workgroups share one system, each owning a strip of cells along the first
 * lattice dimension plus the halo one window (or_sweeps overrelaxation sweeps
 * and a Metropolis sweep) needs.  No counterpart in the reference: a host-side
 * inspection of how csmc_pt_advance would split a system.  stats receives
 * [0] parts, [1] colour phases per window, [2] most local sites of a part,
 * [3] most register slots of a part, [4] the largest part's wave updates per
 * window, [5] the unsplit system's, [6] violations found by replaying the
 * window with a validity bit per site (0 for a sound plan), then per part
 * [7 + 4p] local sites, owned sites, halo sites, boundary sites.  stats must
 * hold 7 + 4 * parts values.  CSMC_EINVAL if the lattice cannot be split so. */
int32_t csmc_split_plan(const csmc_unit_cell *uc, const int32_t *shape, int32_t periodic,
                        int32_t parts, int32_t or_sweeps, int64_t *stats);
/* Host only: the plan of the tiled batch-wide sweeps of a lattice (tiles of
 * cells swept from LDS with partners found by index arithmetic).  Every
 * term of a lattice built from a unit cell joins cells at fixed offsets
 * under the periodic mod of lattice.jl:101-109 (tables of
 * lattice.jl:176-206); a plan exists when the lattice is periodic, has
 * bilinear terms only, and its colouring is lattice-periodic.  head receives
 * CSMC_TILE_HEAD ints: D, nb, K2, p, cz, cinv, ncol, ntiles, L[3], T[3],
 * lo[3], B[3], nt[3], c[3], boxcells, nbox, tcells, ncell, tab_off, nsw
 * (all zero and the reason in csmc_last_error when there is no plan); tab
 * receives tab_len ints e[nb], runbase[p nb], delta[nb K2],
 * rect[nsw ncol nb 6], rawc[ncol] when tab_cap is large enough.  prec: 32
 * or 64 (the record size in LDS); lds_budget: LDS bytes a tile may use;
 * sweeps: 1..8 sweeps run from one staging of a tile's box (the box then
 * holds the partners of the partners, sweeps deep). */
#define CSMC_TILE_HEAD 32
int32_t csmc_tile_plan(const csmc_unit_cell *uc, const int32_t *shape, int32_t periodic, int32_t prec,
                       int64_t lds_budget, int32_t sweeps, int32_t *head, int32_t *tab, int32_t tab_cap,
                       int32_t *tab_len);
/* Lattice(shape, uc, S; bc) (lattice.jl:65-291): builds the per-site tables,
 * colours the interaction hypergraph and uploads device tables.
 * periodic: 1 = "periodic", 0 = "open", other = CSMC_EBC. */
int32_t csmc_lattice_create(csmc_ctx *ctx, const csmc_unit_cell *uc, const int32_t *shape,
                            int32_t periodic, double S, int32_t *lattice_id);
int32_t csmc_lattice_info_get(csmc_ctx *ctx, int32_t lattice_id, csmc_lattice_info *info);
/* The reference's Lattice fields (lattice.jl:12-22) in reference layout:
 * pos[N][D] (site_positions transposed), field[N][3], onsite[N][9],
 * bsite[N][n2], bmat[N][n2][9], csite[N][n3][2], ctens[N][n3][27],
 * qsite[N][n4][3], qtens[N][n4][81].  Any pointer may be NULL. */
int32_t csmc_lattice_tables(csmc_ctx *ctx, int32_t lattice_id, double *pos, double *field,
                            double *onsite, int64_t *bsite, double *bmat, int64_t *csite,
                            double *ctens, int64_t *qsite, double *qtens);
/* no reference counterpart: colour of each site (reference order), 0-based */
int32_t csmc_lattice_colouring(csmc_ctx *ctx, int32_t lattice_id, int32_t *colour);
/* no reference counterpart: replace the couplings of a lattice whose unit
 * cell has the same terms (same bases, offsets, order) -- used to fit
 * couplings without rebuilding neighbour tables */
int32_t csmc_lattice_update_couplings(csmc_ctx *ctx, int32_t lattice_id,
                                      const csmc_unit_cell *uc);

/* ---- batches of systems (replicas or independent lattices) --------------- */
/* One MonteCarlo per system (metropolis.jl:15-29); the batch is what replaces
 * one MPI rank per replica (monte_carlo.jl:246-262). */
int32_t csmc_batch_create(csmc_ctx *ctx, int32_t n_systems, const int32_t *lattice_ids,
                          int32_t precision, int32_t *batch_id);
int32_t csmc_batch_destroy(csmc_ctx *ctx, int32_t batch_id);
/* Lattice.spins get / set (lattice.jl:293-301) for one system, double[N][3] */
int32_t csmc_set_spins(csmc_ctx *ctx, int32_t batch_id, int32_t system, const double *spins);
int32_t csmc_get_spins(csmc_ctx *ctx, int32_t batch_id, int32_t system, double *spins);
/* random_spin_orientation (lattice.jl:306-311) for every site of every system
 * from the Philox stream (seed); initialCondition :random (lattice.jl:76-79) */
int32_t csmc_randomize_spins(csmc_ctx *ctx, int32_t batch_id, uint64_t seed);

/* ---- Hamiltonian (hamiltonian.jl) ----------------------------------------- */
/* get_local_field (hamiltonian.jl:3-67) at every site of one system: H[N][3] */
int32_t csmc_local_field(csmc_ctx *ctx, int32_t batch_id, int32_t system, double *H);
/* energy(lattice, point) (hamiltonian.jl:139-196) at every site: e[N] */
int32_t csmc_site_energy(csmc_ctx *ctx, int32_t batch_id, int32_t system, double *e);
/* get_local_field(lattice, point) and energy(lattice, point)
 * (hamiltonian.jl:3-67,139-196) at a few sites, in time independent of the
 * lattice size: the spins of n_set sites (reference indices set_sites,
 * 0-based, values set_spins[n_set][3]) are first stored into the system,
 * then the FP64 local field H[n_at][3] and site energy e[n_at] (either may
 * be null) of the sites at_sites are evaluated.  A caller whose host copy
 * of the spins changes between calls passes each evaluated site and its
 * partners in set_sites, so the device copy is current where it is read.
 * Counts up to CSMC_SITE_EVAL_MAX; CSMC_EBOUNDS for a site out of range. */
#define CSMC_SITE_EVAL_MAX 65536
int32_t csmc_site_eval(csmc_ctx *ctx, int32_t batch_id, int32_t system, int32_t n_set, const int64_t *set_sites,
                       const double *set_spins, int32_t n_at, const int64_t *at_sites, double *H, double *e);
/* total_energy (hamiltonian.jl:70-132), FP64 reduction, per system: E[n_sys];
 * parts[n_sys][5] (optional) = bilinear, cubic, quartic, zeeman, onsite sums
 * before the 1/2, 1/3, 1/4 weights of hamiltonian.jl:131 */
int32_t csmc_total_energy(csmc_ctx *ctx, int32_t batch_id, double *E, double *parts);
/* get_magnetization (observables.jl:12-18) as the vector sum, FP64: M[n_sys][3] */
int32_t csmc_magnetization(csmc_ctx *ctx, int32_t batch_id, double *M);

/* ---- sweeps ------------------------------------------------------------- */
/* n_sweeps coloured sweeps of `kind` on every system, each system at its
 * temperature T[sys] (sigma[sys] cone width for the cone kinds; adaptive
 * kinds update sigma in place, metropolis.jl:129-131).  accepted[sys]
 * (optional) receives the accepted-proposal count summed over the sweeps.
 * Replaces one call of the alg plugin per sweep (metropolis.jl:181-199),
 * overrelaxation! (monte_carlo.jl:126-139) and the deterministic loop
 * (monte_carlo.jl:201-213).  Philox key = seed; the per-system Metropolis
 * sweep counter advances by one per Metropolis sweep. */
int32_t csmc_sweep(csmc_ctx *ctx, int32_t batch_id, int32_t kind, const double *T,
                   double *sigma, int32_t n_sweeps, uint64_t seed, double *accepted);
/* The plain unbatched form of csmc_sweep: one system at a time, one kernel
 * launch per colour pass, no persistence and no graph.  Same results. */
int32_t csmc_sweep_plain(csmc_ctx *ctx, int32_t batch_id, int32_t kind, const double *T,
                         double *sigma, int32_t n_sweeps, uint64_t seed, double *accepted);
/* Set / read the per-system Metropolis sweep counters (Philox counter word). */
int32_t csmc_set_counters(csmc_ctx *ctx, int32_t batch_id, const uint32_t *counters);
int32_t csmc_get_counters(csmc_ctx *ctx, int32_t batch_id, uint32_t *counters);
/* no reference counterpart: the storage mode the persistent sweep kernel
 * uses for this batch (3 register tables + LDS spins, 2 LDS spins and
 * neighbour tables, 1 LDS spins, 0 global memory); the environment variable
 * CSMC_LDS=0..3, read at context creation, caps it. */
int32_t csmc_batch_storage_mode(csmc_ctx *ctx, int32_t batch_id, int32_t *mode);

/* ---- metropolis_constraint_adaptive! (metropolis.jl:155-179) ---------------
 * The reference evaluates mc.constraint(mc.lattice), any Julia function, per
 * proposal.  The device form takes a constraint that is a sum over sites of a
 * linear and a quadratic term per basis:
 *     c = sum_i (a_b . s_i + s_i . Q_b s_i) - c0,   b = basis of site i,
 * a[n_basis][3], Q[n_basis][9] row-major (bases past n_basis contribute 0),
 * with weight = MonteCarlo.weight and sigma0 = MonteCarlo.sigma0 (the cone
 * width the adaptive law resets to).  Each proposal is a cone move (width
 * sigma[sys]); c is evaluated with the proposed spin in place; new_lambda =
 * lambda - weight * c; acceptance on dE - new_lambda * c in FP64; lambda
 * moves on with every accepted proposal.  Sites are visited colour class by
 * colour class in the device order (the per-proposal chain of lambda runs in
 * that order, one thread per system).  After each sweep sigma follows
 * metropolis.jl:176-177.  Sets or replaces the constraint of every system of
 * the batch; lambda starts at 0 per system (MonteCarlo(), monte_carlo.jl:76)
 * the first time.  Then csmc_sweep / csmc_anneal with
 * kind = CSMC_CONSTRAINT_ADAPTIVE run it (csmc_pt_init refuses it). */
int32_t csmc_constraint_set(csmc_ctx *ctx, int32_t batch_id, int32_t n_basis, const double *a,
                            const double *Q, double c0, double weight, double sigma0);
/* MonteCarlo.lambda per system: set (may be NULL) then get (may be NULL) */
int32_t csmc_constraint_lambda(csmc_ctx *ctx, int32_t batch_id, const double *set, double *get);

/* ---- simulated_annealing! (monte_carlo.jl:157-190) ------------------------ */
/* For each temperature temps[k] (k < n_temps, the caller evaluates the
 * schedule: T0, schedule(1), ... while T > mc.T), every system runs
 * t_thermalization-1 sweep steps: overrelaxation every step (unless
 * or_rate == 0) and `alg` every or_rate-th step (every step if 0); sigma is
 * reset to sigma0 per temperature.  accepted[sys][k] receives the accepted
 * count per temperature.  All systems anneal in one persistent launch. */
int32_t csmc_anneal(csmc_ctx *ctx, int32_t batch_id, const double *temps, int32_t n_temps,
                    int64_t t_thermalization, int64_t overrelaxation_rate, int32_t alg,
                    double sigma0, uint64_t seed, double *accepted);

/* ---- parallel_tempering! (monte_carlo.jl:235-398) ------------------------- */
/* The batch's systems are the replicas; temps[n_sys] ascending-by-slot
 * temperature ladder (slot r plays MPI rank r).  Exchange = temperature-label
 * swap on device, decided in FP64. */
int32_t csmc_pt_init(csmc_ctx *ctx, int32_t batch_id, const double *temps,
                     const csmc_pt_params *params, double sigma0, uint64_t seed);
/* Advance the schedule from the current sweep up to (not including) sweep
 * `until` (or to the end if until < 0).  The call stops early at sweeps where
 * the reference performs host work (checkpoint_rate, report_interval);
 * *stopped receives the sweep it stopped at and *event the reason
 * (0 = until reached, 1 = checkpoint due, 2 = report due, 3 = finished). */
int32_t csmc_pt_advance(csmc_ctx *ctx, int32_t batch_id, int64_t until, int64_t *stopped,
                        int32_t *event);
/* Per temperature slot: slot->system map, FP64 energy of the configuration
 * at that slot, local accepted count, exchange accepted count, sigma. */
int32_t csmc_pt_state(csmc_ctx *ctx, int32_t batch_id, int32_t *system_of_slot, double *E,
                      double *accepted_local, double *accepted_exchange, double *sigma);
/* Workgroups per system of the chained PT kernel chosen at csmc_pt_init:
 * 1 = one workgroup per system; G > 1 = either each system split into G
 * strips with halos (lattices held in registers and LDS, when the environment
 * variable CSMC_SPLIT is "auto" or 2..8), or, for lattices whose spins stay in
 * device memory, G workgroups sharing each colour class of a system with a
 * barrier per colour phase (chosen from the CU and system counts, or forced
 * by CSMC_PARTS0=1..8).  No counterpart in the reference (its replicas are
 * MPI ranks). */
int32_t csmc_pt_split_parts(csmc_ctx *ctx, int32_t batch_id, int32_t *parts);
/* The plain form over the whole batch: the Philox keys and counters of
 * csmc_sweep_plain, so the results are the same bit for bit.  When every
 * lattice of the batch has a tile plan (csmc_batch_tiles), one launch per
 * sweep, or per several sweeps (CSMC_TILE_SWEEPS, default 2, where every
 * lattice has such a plan), sweeps LDS tiles of every system (index-free
 * partners, one read and one write of each spin record per launch);
 * otherwise one launch per colour pass covers every system (one thread per
 * site).  For systems too large
 * for one workgroup to sweep well; not for the adaptive cone.  With
 * csmc_timing on, events bracket the sweeps. */
int32_t csmc_sweep_passes(csmc_ctx *ctx, int32_t batch_id, int32_t kind, const double *T, const double *sigma,
                          int32_t n_sweeps, uint64_t seed, double *accepted);
/* Tiles per system of the batch-wide sweeps (csmc_sweep_passes): > 0 when
 * every lattice of the batch has a tile plan (csmc_tile_plan), so that a
 * sweep runs as one launch of LDS tiles reading one spin buffer and writing
 * the other; 0 when the sweeps run as one launch per colour pass (the reason
 * in csmc_last_error).  sweeps (may be null) receives the sweeps a tiled
 * launch runs while at least that many remain (1 when tiles are off).  No
 * reference counterpart: a query of this engine's schedule. */
int32_t csmc_batch_tiles(csmc_ctx *ctx, int32_t batch_id, int32_t *tiles, int32_t *sweeps);

/* Autotuner (no reference counterpart): time n_sweeps sweeps of `kind` (T
 * for the Metropolis kinds) from the batch's current state in each form --
 * the persistent kernel capped at storage mode 3, 2, 1, 0 (caps that select
 * a mode already timed are skipped) and the batch-wide colour passes -- and
 * keep the fastest for this batch's csmc_sweep calls.  The state (spins and
 * Philox counters) is restored afterwards.  times[5] receives each form's
 * best of two runs in ms (NaN where skipped), *choice its index.  The forms
 * compute the same schedule with the same draws: storage modes agree bit for
 * bit, the passes to rounding. */
int32_t csmc_autotune(csmc_ctx *ctx, int32_t batch_id, int32_t kind, const double *T, int32_t n_sweeps,
                      double *times, int32_t *choice);

/* ---- graph capture (no reference counterpart: launch-bound loops replayed
 * as one hipGraph) ------------------------------------------------------------ */
/* Capture n_sweeps sweeps of `kind` of the plain form (csmc_sweep_plain: one
 * launch per colour pass per system) into an executable hipGraph.  T and
 * sigma (null: 60) are copied to the batch now and read by every replay; the
 * Philox sweep counters are read and advanced on the device, so each replay
 * continues the random stream as the eager calls would, and the adaptive
 * cone width is updated on the device (metropolis.jl:129-131).  The batch
 * must outlive the graph. */
int32_t csmc_graph_sweeps(csmc_ctx *ctx, int32_t batch_id, int32_t kind, const double *T, const double *sigma,
                          int32_t n_sweeps, uint64_t seed, csmc_graph **out);
/* Enqueue one replay on the context's stream (returns without waiting). */
int32_t csmc_graph_launch(csmc_ctx *ctx, csmc_graph *graph);
/* Accepted proposals per system of the last replay, and the cone widths
 * (either may be null); waits for the stream. */
int32_t csmc_graph_accepted(csmc_ctx *ctx, csmc_graph *graph, double *accepted, double *sigma);
int32_t csmc_graph_destroy(csmc_graph *graph);

/* ---- resume (no reference counterpart: the reference's checkpoints,
 * write_MC_checkpoint hdf5.jl:176-180, hold the spins only and cannot
 * restart a run) ---------------------------------------------------------- */
/* The whole state of a parallel-tempering run between two csmc_pt_advance
 * calls (spins, Metropolis counters, slot maps, energies, cone widths,
 * acceptance sums, binners, probe series, exchange records, correlation sums
 * and the sweep position) as one opaque buffer.  buf = NULL returns the size
 * in *bytes; otherwise *bytes is the buffer's capacity on entry and the
 * length written on return. */
int32_t csmc_pt_save(csmc_ctx *ctx, int32_t batch_id, void *buf, int64_t *bytes);
/* The same state taken without waiting: its copy to page-locked host memory
 * is queued on the context's stream after the launches so far (later
 * csmc_pt_advance launches queue behind it) and the call returns at once.
 * csmc_pt_save_wait waits for that copy and returns the buffer (buf = NULL:
 * its size); one save at a time per batch. */
int32_t csmc_pt_save_async(csmc_ctx *ctx, int32_t batch_id);
int32_t csmc_pt_save_wait(csmc_ctx *ctx, int32_t batch_id, void *buf, int64_t *bytes);
/* Continue a saved run: after csmc_pt_init (and csmc_pt_set_correlations)
 * with the temperatures, parameters, seed and momenta of the saved run on a
 * batch of the same lattice, size and precision, restore the state; the
 * following csmc_pt_advance calls continue the run as if it had never
 * stopped.  A state that does not match fails with CSMC_EINVAL. */
int32_t csmc_pt_load(csmc_ctx *ctx, int32_t batch_id, const void *buf, int64_t bytes);

/* Final observables per slot (hdf5.jl:204-239 / observables.jl:36-63):
 * out[n_sys][8] = C, dC, chi, dchi, <E>, dE, <|M|>, d|M| (absolute errors). */
int32_t csmc_pt_observables(csmc_ctx *ctx, int32_t batch_id, double *out);

/* compute_equal_time_correlations(lat, ks) (spin_correlations.jl:6-43) for
 * every system of the batch (which must share one lattice): ks is the
 * reference's D x n_k momentum matrix (double[n_k][D]); out receives, per
 * system, its 9 x n_k Suv matrix as Julia lays it out (double[n_sys][n_k][9]:
 * xx xy xz yx yy yz zx zy zz per momentum), already divided by lat.size.  FP64
 * throughout: one fused MFMA kernel (phases in registers), or with
 * CSMC_SSF=blas set at csmc_ctx_create, phase matrices and rocBLAS GEMMs. */
int32_t csmc_spin_correlations(csmc_ctx *ctx, int32_t batch_id, const double *ks, int32_t n_k,
                               double *out);

/* MonteCarlo(...; corr = true, ks) for parallel tempering
 * (monte_carlo.jl:61-81,371-375): after csmc_pt_init and before the first
 * csmc_pt_advance, measure the equal-time correlations at ks (double[n_k][D])
 * at every probe sweep, after that sweep's exchange, and add them to the sum
 * of the temperature slot each system holds.  n_k = 0 turns it off. */
int32_t csmc_pt_set_correlations(csmc_ctx *ctx, int32_t batch_id, const double *ks, int32_t n_k);
/* mean(mc.observables.correlations) per temperature slot (hdf5.jl:229-233):
 * mean[n_sys][n_k][9] (each slot's 9 x n_k Suv in Julia's layout; NaN before
 * the first probe) and the number of probes. */
int32_t csmc_pt_correlations(csmc_ctx *ctx, int32_t batch_id, double *mean, int64_t *count);

/* ---- differentiation (no reference counterpart: adjoint of
 * total_energy, hamiltonian.jl:70-132) ------------------------------------ */
/* dE/dS_i for every site of every system, double[n_sys][N][3]. */
int32_t csmc_energy_grad_spins(csmc_ctx *ctx, int32_t batch_id, double *dEdS);
/* dE/d(coupling) per unit-cell term, per system:
 * out[n_sys][n_field*3 + n_onsite*9 + n2*9 + n3*27 + n4*81] in unit-cell term
 * order, each tensor in the csmc_unit_cell layout. */
int32_t csmc_energy_grad_couplings(csmc_ctx *ctx, int32_t batch_id, double *dEdJ);
/* No counterpart in the reference.  csmc_energy_grad_couplings for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * fitting loops, harnesses that forbid hipMalloc): the same kernel and
 * output, with the per-system gradients staged in a fixed device array of
 * CSMC_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead
 * of a buffer allocated per call.  No hipMalloc / hipFree or allocating
 * library routine runs.  Limits: n_sys * (gradient length) * 8 bytes must not
 * exceed CSMC_NOALLOC_SCRATCH_BYTES (CSMC_EINVAL otherwise), and the
 * lattice tables must already be on the device (any csmc_set_spins,
 * csmc_randomize_spins or sweep uploads them; CSMC_ESTATE otherwise).  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (from several contexts or threads) race. */
#define CSMC_NOALLOC_SCRATCH_BYTES (1 << 20)
int32_t csmc_energy_grad_couplings_noalloc(csmc_ctx *ctx, int32_t batch_id, double *dEdJ);
/* One projected gradient-descent step on the sphere for every system:
 * s <- S * normalise(s - eta * (g - (g.s^) s^)), g = dE/ds. */
int32_t csmc_gradient_descent(csmc_ctx *ctx, int32_t batch_id, double eta, int32_t n_steps);

#ifdef __cplusplus
}
#endif
#endif
