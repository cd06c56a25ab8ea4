// api.cpp -- libcsmc host runtime and C ABI (include/csmc.h).
//
// A context owns one HIP stream on one device, the device table pools of
// every lattice created on it, and the batches of systems.  Lattice tables
// are append-only, so pool offsets stay valid when lattices are added; the
// device pools are re-uploaded lazily before the next launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/csmc.h"
#include "csmc_internal.h"
#include "kernels.h"
#include "lattice.h"
#include "split.h"
#include "tile.h"

using namespace csmc;

// Device scratch of the no-allocation entry point
// (csmc_energy_grad_couplings_noalloc): one array per GPU, held until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: host code uses only the address hipGetSymbolAddress gives
// for the current device, kept per device by scratch_setup().
static __device__ __align__(256) unsigned char scratch[CSMC_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device -> address of `scratch` there

// once per device, with that device current (csmc_ctx_create)
int scratch_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return CSMC_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(CSMC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return CSMC_OK;
}

void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    return it == g_scratch_ptr.end() ? nullptr : it->second;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(CSMC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void free_() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // keeps the buffer when it already holds count elements
    hipError_t reserve(size_t count) {
        if (p && n >= count) return hipSuccess;
        return alloc(count);
    }
    hipError_t alloc(size_t count) {
        free_();
        n = count;
        if (!count) return hipSuccess;
        return hipMalloc((void **)&p, count * sizeof(T));
    }
    ~DevBuf() { free_(); }
};

// page-locked host memory (hipHostMalloc), kept until the owner goes
struct Pinned {
    unsigned char *p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t count) {
        if (p && n >= count) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc((void **)&p, count ? count : 1, hipHostMallocDefault);
        if (e == hipSuccess) n = count;
        return e;
    }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

// the inputs of a correlation call on the device: positions (in the order
// the kernels visit the sites), the momenta and the device site of each row
struct SsfBufs {
    DevBuf<double> pos, ks;
    DevBuf<int32_t> perm, tiles;
};

struct PTHost {
    bool init = false;
    csmc_pt_params p{};
    int64_t sweep = 0;
    bool mid = false;
    int64_t total = 0, dosweep = 1;
    std::vector<double> temps;
    DevBuf<double> temps_d, E, sigma, acc_local, acc_exch, bins;
    DevBuf<int32_t> slot_of_sys, sys_of_slot;
    DevBuf<uint32_t> done;
    // chained launches: the exchange records ([PT_RECBUF][R][4]) and the
    // wait-failure flag
    DevBuf<uint64_t> rec;
    DevBuf<uint32_t> err;
    bool chain_ok = true;  // cleared when a cooperative launch is refused
    // probe series not yet folded into the binners: probes [0, probe_pos)
    DevBuf<double> series;
    int64_t probe_pos = 0;
    uint64_t seed = 0;
    // split form of the chained kernel: split_G workgroups per system (1:
    // the whole-system kernel), its tables, exchange words and epochs
    int split_G = 1, split_nor = 0, split_spt = 0, split_bmax = 0;
    size_t split_lds = 0;
    DevBuf<int32_t> split_tab;
    DevBuf<uint64_t> split_hx, split_px;
    DevBuf<uint32_t> split_epoch;
    // chained kernel in storage mode 0: parts0 workgroups per system (1: one),
    // their barriers ([R][2]) and workgroup sums ([R][2][parts0][8])
    int parts0 = 1;
    int parts_lm = -1;         // the storage mode parts0 was chosen for
    bool parts_ready = false;  // parts0 chosen (at csmc_pt_init or the first
                               // persistent launch in storage mode 0)
    DevBuf<uint32_t> bar0, err0, xcc0;
    DevBuf<double> psum0;
    // equal-time correlations at every probe (corr = true): the sum over
    // probes per temperature slot, [R][corr_nk][9], and the probe count
    int corr_nk = 0;
    int64_t corr_count = 0;
    SsfPlan corr_plan{};
    SsfBufs corr_in;
    DevBuf<double> corr_acc;
    // asynchronous save: the state copied to page-locked host memory in
    // stream order; the host scalars of that moment in snap_head
    Pinned snap_host;
    std::vector<unsigned char> snap_head;
    size_t snap_bytes = 0;
    hipEvent_t snap_done = nullptr;
    bool snap_pending = false;
    ~PTHost() {
        if (snap_done) (void)hipEventDestroy(snap_done);
    }
};

constexpr int PROBE_CAP = 4096;  // probes per slot held before a fold

struct Batch {
    int prec = 64;
    int nsys = 0;
    int maxN = 0;
    int maxK2 = 0;
    int k2c = 0;  // bilinear slot count shared by all systems (0 if mixed)
    int nslots = 0;  // most register-table slots of any system
    int64_t total = 0;
    bool self = false;
    bool gen = false;  // some lattice has cubic or quartic terms
    std::vector<int> lat;
    std::vector<SysDev> sys;
    DevBuf<SysDev> sys_d;
    DevBuf<unsigned char> spins;
    DevBuf<unsigned char> spins_alt;  // the other buffer of csmc_gradient_descent and the tiled sweeps
    DevBuf<uint32_t> mcount;
    DevBuf<double> T, sigma, acc;  // per-system scratch
    DevBuf<uint32_t> accU;
    DevBuf<double> stage;          // host<->device staging (total*3 doubles)
    DevBuf<unsigned char> sites;   // csmc_site_eval: indices and values of a few sites
    // equal-time correlations: device positions, momenta and GEMM workspaces
    SsfBufs ssf_in;
    DevBuf<double> ssf_C, ssf_Q, ssf_B, ssf_Cre, ssf_Cim, ssf_P, ssf_out;
    // set by csmc_autotune: the highest storage mode the persistent kernel
    // may use, and whether csmc_sweep runs as batch-wide colour passes
    int lm_cap = 3;
    bool passes = false;
    PTHost pt;
    // packed schedule of sweeps and annealing in storage mode 0 for batches
    // of uneven lattices (pack_plan)
    struct Pack {
        int lm = -1, cap = -1;   // made for this storage mode and cap
        bool use = false;
        int nwg = 0, gmax = 1;
        double makespan = 0, naive = 0;
        DevBuf<int32_t> off;
        DevBuf<int4> tasks;
        DevBuf<uint32_t> bar0, xcc0;
        DevBuf<double> psum0;
    } pack;
    // metropolis_constraint_adaptive! (csmc_constraint_set): the per-basis
    // coefficients, the multiplier lambda, constraint value and accepted
    // counts per system, and per spin slot the proposal, energy change,
    // constraint change, uniform and decision of the current colour pass
    bool cons = false;
    int cons_nb = 0;
    double cons_c0 = 0.0, cons_w = 0.0, cons_sigma0 = 60.0;
    DevBuf<double> cons_coef, cons_lambda, cons_c, cons_acc, cons_total, cons_dE, cons_dC, cons_u;
    DevBuf<unsigned char> cons_prop, cons_dec;
    // tiled sweeps (tile.cpp / tile.hip) of the batch-wide sweep form:
    // state 0 not planned yet, 1 every lattice has a plan, -1 some has none
    struct TileSet {
        size_t lds = 0;
        int max_tiles = 0;
        bool flat = true;  // every plan's first position has extent 1
        DevBuf<TileDev> td;
        DevBuf<int32_t> tab;
    };
    struct Tiles {
        int state = 0;
        std::string why;
        TileSet one;    // one sweep per launch
        int sweeps = 1;  // > 1: every lattice also has a plan of this many
        TileSet many;   // sweeps per launch
    } tiles;
    size_t spin_bytes() const { return prec == 32 ? 12 : 24; }  // packed {x,y,z} records
};

}  // namespace

struct csmc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    csmc_ctx *ctx = nullptr;
    int32_t batch = -1;
    int32_t kind = 0;
    void *spins = nullptr;  // the spin buffer the recorded launches use
};

struct csmc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<HostLattice>> lats;
    std::vector<LatDev> h_lat;
    std::vector<int32_t> h_i32;
    std::vector<int4> h_runs;
    std::vector<double> h_coef;
    DevBuf<LatDev> d_lat;
    DevBuf<int32_t> d_i32;
    DevBuf<int4> d_runs;
    DevBuf<float> d_coefF;
    DevBuf<double> d_coefD;
    bool dirty = true;
    std::map<int, std::unique_ptr<Batch>> batches;
    int next_batch = 1;
    int max_lds = 65536;
    int lds_cap = 3;  // highest workgroup storage mode allowed
    bool specialise = true;  // use slot-count specialised kernels when they exist
    bool pt_chain = true;    // chain PT exchange periods in cooperative launches
    int split = 1;           // parts per system of the split PT kernel: 1 never
                             // split (default), 0 chosen from the plan's cost,
                             // 2..8 forced
    int ncu = 0;             // compute units of the device
    void *blas = nullptr;    // rocBLAS handle (created on first use)
    bool ssf_blas = false;   // equal-time correlations through rocBLAS
                             // instead of the fused MFMA kernel
    bool dense_tensors = false;  // contract every cubic / quartic coupling in
                                 // full (no sparsity patterns)
    bool verbose = false;        // CSMC_VERBOSE=1: device and lattice notes on stderr
    bool pack = true;            // packed schedules for uneven batches (CSMC_PACK=0: off)
    bool tile = true;            // tiled batch-wide sweeps where a lattice allows (CSMC_TILE=0: off)
    bool parttab = true;         // storage mode 0: partners of each workgroup's sites in LDS
                                 // (CSMC_PARTTAB=0: off)
    size_t tile_lds = 48 << 10;  // LDS budget of one tile (CSMC_TILE_LDS, KiB)
    // ... of a tile of one sweep a launch (CSMC_TILE_LDS1, KiB): larger
    // tiles, fewer workgroups per CU; on the honeycomb L=512 probe the
    // Metropolis sweep took 568 us with 32 x 64-cell tiles (64 KB) against
    // 620 with 32 x 32 (48 KB) and 1130 with 16 x 32 (24 KB)
    size_t tile_lds1 = 64 << 10;
    int tile_sweeps = 2;         // sweeps per tiled launch where a plan allows (CSMC_TILE_SWEEPS, 1..8)
    int parts0 = 0;              // workgroups per system of the chained PT kernel in
                                 // storage mode 0: 0 chosen, 1 off, 2..8 forced
    uint32_t spin_limit = 1u << 26;  // polls of a wait for another workgroup
                                     // (x 1/16 for the exchange waits)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
};

namespace {

// no cubic / quartic slot of the batch's lattices is contracted in full
bool batch_sparse(csmc_ctx *c, Batch *b) {
    for (int id : b->lat)
        if (c->h_lat[id].flags & LAT_DENSE_T) return false;
    return true;
}

// LDS bytes of the per-workgroup partner tables of storage mode 0 (the
// sparse general kernels) with G workgroups per system: every lattice of the
// batch within uint16 indices, the busiest thread's visits x the block size
// x the partner rows; 0 when some lattice does not qualify or the table does
// not fit beside the reduction scratch
size_t part_table_bytes(csmc_ctx *c, Batch *b, int G) {
    if (!c->parttab || b->prec != 32 || !(b->self || b->gen) || !batch_sparse(c, b)) return 0;
    const int threads = wg_threads_for(b->prec, true, true);
    size_t most = 0;
    for (int id : b->lat) {
        const HostLattice &L = *c->lats[id];
        if (L.N > 65535) return 0;
        size_t nvisit = 0;
        for (const auto &r : L.runs) nvisit += (size_t)(r[1] - r[0] + G * threads - 1) / (G * threads);
        most = std::max(most, nvisit * threads * (size_t)(L.K2 + 2 * L.K3 + 3 * L.K4) * 2);
    }
    return RED_BYTES + most <= (size_t)c->max_lds ? most : 0;
}


int ensure_pools(csmc_ctx *c) {
    if (!c->dirty) return CSMC_OK;
    HIPCHK(c->d_lat.alloc(c->h_lat.size()));
    HIPCHK(c->d_i32.alloc(c->h_i32.size()));
    HIPCHK(c->d_runs.alloc(c->h_runs.size()));
    HIPCHK(c->d_coefF.alloc(c->h_coef.size()));
    HIPCHK(c->d_coefD.alloc(c->h_coef.size()));
    std::vector<float> cf(c->h_coef.begin(), c->h_coef.end());
    if (c->h_lat.size())
        HIPCHK(hipMemcpy(c->d_lat.p, c->h_lat.data(), c->h_lat.size() * sizeof(LatDev), hipMemcpyHostToDevice));
    if (c->h_i32.size())
        HIPCHK(hipMemcpy(c->d_i32.p, c->h_i32.data(), c->h_i32.size() * 4, hipMemcpyHostToDevice));
    if (c->h_runs.size())
        HIPCHK(hipMemcpy(c->d_runs.p, c->h_runs.data(), c->h_runs.size() * sizeof(int4), hipMemcpyHostToDevice));
    if (cf.size()) {
        HIPCHK(hipMemcpy(c->d_coefF.p, cf.data(), cf.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_coefD.p, c->h_coef.data(), c->h_coef.size() * 8, hipMemcpyHostToDevice));
    }
    c->dirty = false;
    return CSMC_OK;
}

DevPools pools(csmc_ctx *c, Batch *b) {
    DevPools P;
    P.lats = c->d_lat.p;
    P.i32 = c->d_i32.p;
    P.runs = c->d_runs.p;
    P.coefF = c->d_coefF.p;
    P.coefD = c->d_coefD.p;
    P.sys = b ? b->sys_d.p : nullptr;
    return P;
}

int pad4(int k) { return (k + 3) & ~3; }

// device coefficient block of a lattice: the host block with the bilinear
// slots padded to a multiple of 4 (zero couplings in the padding)
std::vector<double> device_coef(const HostLattice &L) {
    const int K2P = pad4(L.K2);
    std::vector<double> v(L.coef.begin(), L.coef.begin() + L.coef_c2);
    for (int b = 0; b < L.nb; ++b) {
        const double *src = &L.coef[L.coef_c2 + (size_t)b * L.K2 * 9];
        v.insert(v.end(), src, src + (size_t)L.K2 * 9);
        v.insert(v.end(), (size_t)(K2P - L.K2) * 9, 0.0);
    }
    v.insert(v.end(), L.coef.begin() + L.coef_c3, L.coef.end());
    return v;
}

// LAT_* flags of a lattice's couplings: which optional terms are non-zero,
// and whether every bilinear matrix is diagonal
int coupling_flags(const HostLattice &L) {
    int flags = LAT_DIAG;
    for (long k = 0; k < (long)L.nb * 9; ++k)
        if (L.coef[L.coef_onsite + k] != 0.0) flags |= LAT_ONSITE;
    for (long k = 0; k < (long)L.nb * 3; ++k)
        if (L.coef[L.coef_field + k] != 0.0) flags |= LAT_FIELD;
    for (long m = 0; m < (long)L.nb * L.K2; ++m)
        for (int e = 0; e < 9; ++e)
            if (e % 4 != 0 && L.coef[L.coef_c2 + m * 9 + e] != 0.0) flags &= ~LAT_DIAG;
    return flags;
}

// LAT_GRAD_LF: the energy's gradient equals the local field.  With the
// 1/2 of hamiltonian.jl:131, dE/ds_x = 1/2 sum_k M_k s_j (x's own bonds) +
// 1/2 sum M^T s_i (bonds i -> x); when every bond i -> j with M has a bond
// j -> i with M^T (as multisets per pair of sites) the second sum equals
// the first, and (O + O^T) s = 2 O s for symmetric O.
bool gradient_is_local_field(const HostLattice &L) {
    if (L.K3 > 0 || L.K4 > 0 || L.has_self) return false;
    for (int b = 0; b < L.nb; ++b)
        for (int u = 0; u < 3; ++u)
            for (int v = u + 1; v < 3; ++v)
                if (L.coef[L.coef_onsite + b * 9 + u * 3 + v] != L.coef[L.coef_onsite + b * 9 + v * 3 + u])
                    return false;
    const int64_t N = L.N;
    auto mat = [&](int64_t ref, int k) { return &L.coef[L.coef_c2 + ((size_t)L.basis_of((int)ref) * L.K2 + k) * 9]; };
    for (int64_t i = 0; i < N; ++i)
        for (int k = 0; k < L.K2; ++k) {
            const int64_t j = L.nbr2[(size_t)i * L.K2 + k];
            if (j < 0) continue;
            const double *M = mat(i, k);
            // bonds i -> j with the same matrix, and bonds j -> i with M^T
            int same = 0, back = 0;
            for (int q = 0; q < L.K2; ++q) {
                if (L.nbr2[(size_t)i * L.K2 + q] == j && std::equal(M, M + 9, mat(i, q))) ++same;
                if (L.nbr2[(size_t)j * L.K2 + q] == i) {
                    const double *T = mat(j, q);
                    bool t = true;
                    for (int u = 0; u < 3 && t; ++u)
                        for (int v = 0; v < 3 && t; ++v) t = T[u * 3 + v] == M[v * 3 + u];
                    back += t;
                }
            }
            if (same != back) return false;
        }
    return true;
}

// Sparsity pattern (TP_*) of every cubic coupling [nb][K3], then of every
// quartic one [nb][K4].  Entry e of a cubic tensor is p + 3q + 9r, of a
// quartic one p + 3q + 9r + 27t (p the site's own component).  The first
// pattern that holds every non-zero entry is taken, so an all-zero slot is
// TP_ZERO; dense = true leaves every non-zero slot TP_DENSE.
std::vector<int32_t> tensor_patterns(const HostLattice &L, bool dense) {
    std::vector<int32_t> out;
    for (long m = 0; m < (long)L.nb * L.K3; ++m) {
        const double *C = &L.coef[L.coef_c3 + m * 27];
        bool zero = true, alt = true;
        for (int e = 0; e < 27; ++e) {
            if (C[e] == 0.0) continue;
            const int p = e % 3, q = e / 3 % 3, r = e / 9;
            zero = false;
            alt = alt && p != q && q != r && p != r;
        }
        out.push_back(zero ? TP_ZERO : (alt && !dense) ? TP_C3_ALT : TP_DENSE);
    }
    for (long m = 0; m < (long)L.nb * L.K4; ++m) {
        const double *Q = &L.coef[L.coef_c4 + m * 81];
        bool zero = true, pq = true, pr = true, pt = true;
        for (int e = 0; e < 81; ++e) {
            if (Q[e] == 0.0) continue;
            const int p = e % 3, q = e / 3 % 3, r = e / 9 % 3, t = e / 27;
            zero = false;
            pq = pq && p == q && r == t;
            pr = pr && p == r && q == t;
            pt = pt && p == t && q == r;
        }
        out.push_back(zero ? TP_ZERO : dense ? TP_DENSE : pq ? TP_Q4_PQ : pr ? TP_Q4_PR : pt ? TP_Q4_PT : TP_DENSE);
    }
    return out;
}

// append one lattice's device tables to the host-side pools
void append_lattice(csmc_ctx *c, const HostLattice &L) {
    LatDev d{};
    const int N = (int)L.N;
    d.N = N;
    d.nb = L.nb;
    d.K2 = L.K2;
    d.K2P = pad4(L.K2);
    d.K3 = L.K3;
    d.K4 = L.K4;
    d.ncol = L.ncol;
    d.nruns = (int)L.runs.size();
    d.has_self = L.has_self;
    d.S = L.S;
    d.flags = coupling_flags(L);
    if (gradient_is_local_field(L)) d.flags |= LAT_GRAD_LF;
    if (L.has_self) {
        bool bc = false;  // a bilinear or cubic slot names its own site
        for (int64_t s = 0; s < L.N && !bc; ++s) {
            for (int k = 0; k < L.K2; ++k) bc = bc || L.nbr2[(size_t)s * L.K2 + k] == s;
            for (int k = 0; k < L.K3 * 2; ++k) bc = bc || L.nbr3[(size_t)s * L.K3 * 2 + k] == s;
        }
        if (!bc) d.flags |= LAT_SELF_Q4;
    }
    auto &I = c->h_i32;
    auto dev = [&](int ref) { return ref < 0 ? -1 : L.dev_of_ref[ref]; };
    auto devp = [&](int ref) { return ref < 0 ? N : L.dev_of_ref[ref]; };  // absent -> pad site
    // neighbour tables in device order, slot-major
    d.nbr_off = (int64_t)I.size();
    for (int k = 0; k < L.K2; ++k)
        for (int i = 0; i < N; ++i) I.push_back(devp(L.nbr2[(size_t)L.ref_of_dev[i] * L.K2 + k]));
    for (int k = L.K2; k < d.K2P; ++k)
        for (int i = 0; i < N; ++i) I.push_back(N);
    for (int k = 0; k < L.K3; ++k)
        for (int q = 0; q < 2; ++q)
            for (int i = 0; i < N; ++i) I.push_back(devp(L.nbr3[((size_t)L.ref_of_dev[i] * L.K3 + k) * 2 + q]));
    for (int k = 0; k < L.K4; ++k)
        for (int q = 0; q < 3; ++q)
            for (int i = 0; i < N; ++i) I.push_back(devp(L.nbr4[((size_t)L.ref_of_dev[i] * L.K4 + k) * 3 + q]));
    // lattices with the same neighbour tables (the same cell structure and
    // shape, whatever the couplings) share one copy: a batch of them then
    // gathers through one table that stays in cache
    for (const LatDev &o : c->h_lat) {
        if (o.N != d.N || o.K2P != d.K2P || o.K3 != d.K3 || o.K4 != d.K4) continue;
        if (std::equal(I.begin() + d.nbr_off, I.end(), I.begin() + o.nbr_off)) {
            I.resize((size_t)d.nbr_off);
            d.nbr_off = o.nbr_off;
            break;
        }
    }
    d.ref_off = (int64_t)I.size();
    for (int i = 0; i < N; ++i) I.push_back(L.ref_of_dev[i]);
    d.sub_off = (int64_t)I.size();
    for (int i = 0; i < N; ++i) I.push_back(L.basis_of(L.ref_of_dev[i]));
    d.colrun_off = (int64_t)I.size();
    for (int v : L.colrun) I.push_back(v);
    d.run_off = (int64_t)c->h_runs.size();
    for (auto &r : L.runs) c->h_runs.push_back(make_int4(r[0], r[1], r[2], r[3]));
    // register-table schedule: each run cut into REG_THREADS-site slots,
    // colour by colour (runs are already in colour order)
    d.slot_off = (int64_t)c->h_runs.size();
    d.colslot_off = (int64_t)I.size();
    int nslots = 0;
    for (int col = 0; col < L.ncol; ++col) {
        I.push_back(nslots);
        for (int r = L.colrun[col]; r < L.colrun[col + 1]; ++r)
            for (int first = L.runs[r][0]; first < L.runs[r][1]; first += REG_THREADS, ++nslots)
                c->h_runs.push_back(make_int4(first, L.runs[r][1], L.runs[r][2], col));
    }
    I.push_back(nslots);
    d.nslots = nslots;
    // basis shared by all runs of a colour (-1 if the colour mixes bases)
    d.flags |= LAT_COLBASIS;
    for (int col = 0; col < L.ncol; ++col) {
        int cb = L.runs[L.colrun[col]][2];
        for (int r = L.colrun[col]; r < L.colrun[col + 1]; ++r)
            if (L.runs[r][2] != cb) cb = -1;
        if (cb < 0) d.flags &= ~LAT_COLBASIS;
        I.push_back(cb);
    }
    if ((d.flags & LAT_COLBASIS) && L.ncol == 2) {
        int size[2] = {0, 0};
        for (int col = 0; col < 2; ++col)
            for (int r = L.colrun[col]; r < L.colrun[col + 1]; ++r) size[col] += L.runs[r][1] - L.runs[r][0];
        const int spc = size[0] / REG_THREADS;
        if (size[0] == size[1] && size[0] % REG_THREADS == 0 && (spc == 1 || spc == 2 || spc == 4))
            d.flags |= LAT_TWO_COLOUR;
    }
    d.pat_off = (int64_t)I.size();
    {
        const std::vector<int32_t> pat = tensor_patterns(L, c->dense_tensors);
        I.insert(I.end(), pat.begin(), pat.end());
        for (int32_t v : pat)
            if (v == TP_DENSE) d.flags |= LAT_DENSE_T;
        if (c->verbose) {
            int n3[5] = {0}, n4[5] = {0};
            for (size_t m = 0; m < pat.size(); ++m) ++(m < (size_t)L.nb * L.K3 ? n3 : n4)[pat[m]];
            fprintf(stderr,
                    "csmc: lattice of %d sites, cubic slots dense %d alt %d zero %d, quartic slots dense %d "
                    "pq %d pr %d pt %d zero %d\n",
                    N, n3[TP_DENSE], n3[TP_C3_ALT], n3[TP_ZERO], n4[TP_DENSE], n4[TP_Q4_PQ], n4[TP_Q4_PR],
                    n4[TP_Q4_PT], n4[TP_ZERO]);
        }
    }
    // incidence lists: participant -> (owner, kind, role, slot)
    std::vector<std::vector<Inc>> inc(N);
    for (int i = 0; i < N; ++i) {
        int rs = L.ref_of_dev[i];
        for (int k = 0; k < L.K2; ++k) {
            int j = L.nbr2[(size_t)rs * L.K2 + k];
            if (j < 0) continue;
            inc[i].push_back({i, 0 | (0 << 2) | (k << 4)});
            inc[dev(j)].push_back({i, 0 | (1 << 2) | (k << 4)});
        }
        for (int k = 0; k < L.K3; ++k) {
            int j = L.nbr3[((size_t)rs * L.K3 + k) * 2];
            if (j < 0) continue;
            inc[i].push_back({i, 1 | (0 << 2) | (k << 4)});
            for (int q = 0; q < 2; ++q)
                inc[dev(L.nbr3[((size_t)rs * L.K3 + k) * 2 + q])].push_back({i, 1 | ((q + 1) << 2) | (k << 4)});
        }
        for (int k = 0; k < L.K4; ++k) {
            int j = L.nbr4[((size_t)rs * L.K4 + k) * 3];
            if (j < 0) continue;
            inc[i].push_back({i, 2 | (0 << 2) | (k << 4)});
            for (int q = 0; q < 3; ++q)
                inc[dev(L.nbr4[((size_t)rs * L.K4 + k) * 3 + q])].push_back({i, 2 | ((q + 1) << 2) | (k << 4)});
        }
    }
    d.inc_off = (int64_t)I.size();
    int64_t acc = 0;
    for (int i = 0; i < N; ++i) {
        I.push_back((int32_t)acc);
        acc += (int64_t)inc[i].size();
    }
    I.push_back((int32_t)acc);
    d.inc_n = acc;
    for (int i = 0; i < N; ++i)
        for (auto &e : inc[i]) {
            I.push_back(e.owner);
            I.push_back(e.code);
        }
    // gradient tables
    d.grad_off = (int64_t)I.size();
    for (int v : L.slot2) I.push_back(v);
    for (int v : L.role2) I.push_back(v);
    for (int v : L.slot3) I.push_back(v);
    for (int v : L.role3) I.push_back(v);
    for (int v : L.slot4) I.push_back(v);
    for (int v : L.role4) I.push_back(v);
    for (int v : L.fsrc) I.push_back(v);
    for (int v : L.osrc) I.push_back(v);
    int nf = (int)L.fb.size(), no = (int)L.ob.size();
    I.push_back(0);
    I.push_back(3 * nf);
    I.push_back(3 * nf + 9 * no);
    I.push_back(3 * nf + 9 * no + 9 * L.n2);
    I.push_back(3 * nf + 9 * no + 9 * L.n2 + 27 * L.n3);
    d.grad_len = L.grad_len;
    d.coef_off = (int64_t)c->h_coef.size();
    std::vector<double> dc = device_coef(L);
    c->h_coef.insert(c->h_coef.end(), dc.begin(), dc.end());
    c->h_lat.push_back(d);
    c->dirty = true;
}

Batch *get_batch(csmc_ctx *c, int id) {
    auto it = c->batches.find(id);
    return it == c->batches.end() ? nullptr : it->second.get();
}

#define CHECK_CTX(c) \
    if (!(c)) return fail(CSMC_EINVAL, "null context")
#define GET_BATCH(c, id, b)                                                     \
    Batch *b = get_batch(c, id);                                                \
    if (!b) return fail(CSMC_ESTATE, "unknown batch id " + std::to_string(id))

// register-table slot budgets compiled (kernels_reg.hip): per thread, SPT
// sites per sweep with K2C neighbour offsets each held in registers (8 slots
// only for up to 4 neighbours); 0 if the batch does not fit
int reg_spt(int nslots, int k2c) {
    if (nslots <= 4) return 4;
    if (nslots <= 8 && k2c <= 4) return 8;
    return 0;
}

// storage mode of the workgroup kernel, capped by CSMC_LDS (0..3) for
// measurements of each mode:
//   3  register tables + float4 spins in LDS (FP32, bilinear-only lattices
//      with a shared slot count of 3, 4 or 6 and at most 8 slots a thread)
//   2  spins as three component arrays + uint16 neighbour tables in LDS
//   1  spins in LDS, tables in global memory
//   0  everything in global memory
int choose_lm(csmc_ctx *c, Batch *b, size_t *bytes) {
    auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
    size_t comp = b->prec == 32 ? 4 : 8;
    size_t spins = al(3 * comp * (size_t)(b->maxN + 1));  // + the zero pad site
    size_t idx = al(2 * (size_t)b->maxK2 * b->maxN);
    size_t avail = (size_t)c->max_lds > RED_BYTES ? (size_t)c->max_lds - RED_BYTES : 0;
    const bool k2ok = b->k2c == 3 || b->k2c == 4 || b->k2c == 6;
    bool colbasis = true;
    for (int id : b->lat) colbasis = colbasis && (c->h_lat[id].flags & LAT_COLBASIS);
    const int cap = std::min(c->lds_cap, b->lm_cap);
    if (cap >= 3 && c->specialise && b->prec == 32 && !b->self && !b->gen && k2ok && colbasis &&
        reg_spt(b->nslots, b->k2c) > 0 && 16 * (size_t)(b->maxN + 1) <= avail) {
        *bytes = 16 * (size_t)(b->maxN + 1);
        return 3;
    }
    if (cap >= 2 && b->maxN < 65535 && spins + idx <= avail) {
        *bytes = spins + idx;
        return 2;
    }
    if (cap >= 1 && spins <= avail) {
        *bytes = spins;
        return 1;
    }
    *bytes = 0;
    return 0;
}

// LDS of a split part: local spins and the pad site, its boundary list, its
// halo pairs
size_t split_lds_bytes(const SplitPlan &p) {
    return 16 * (size_t)(p.maxloc + 1) + 4 * (size_t)p.maxbnd + 8 * (size_t)p.maxhalo + 8;
}

// Split form of the chained PT kernel (kernels_split.hip): plan parts for
// the batch's lattice and upload the tables, or leave split_G = 1.  Off
// unless asked for: measured on the honeycomb L=64 x 64-replica workload the
// whole-system kernel is faster (each colour phase has a fixed latency the
// split does not shorten, and the split adds halo updates and a window
// exchange).  Automatic choice (CSMC_SPLIT=auto): the plan whose largest part
// does the least work, if that is at most 3/4 of the whole system's and
// G * R workgroups fit the device at once.
int setup_split(csmc_ctx *c, Batch *b) {
    PTHost &pt = b->pt;
    pt.split_G = 1;
    const int R = b->nsys;
    size_t whole = 0;
    if (c->split == 1 || !c->pt_chain || R < 2 || choose_lm(c, b, &whole) != 3) return CSMC_OK;
    const HostLattice &L = *c->lats[b->lat[0]];
    if (pt.p.overrelaxation_rate > 64) return CSMC_OK;  // the halo would hold the whole system
    const int nor = pt.p.overrelaxation_rate != 0 ? (int)pt.dosweep : 0;
    const int gmax = c->split > 1 ? c->split : std::min(SPLIT_MAX, c->ncu / R);
    SplitPlan best;
    for (int G = c->split > 1 ? c->split : 2; G <= gmax; ++G) {
        SplitPlan plan;
        if (!build_split(L, G, nor, REG_THREADS, plan)) continue;
        if (RED_BYTES + split_lds_bytes(plan) > (size_t)c->max_lds) continue;
        if (reg_spt(plan.maxslots, b->k2c) == 0) continue;
        if (c->split == 0 && plan.work_max * 4 > plan.work_full * 3) continue;
        if (best.G == 1 || plan.work_max < best.work_max) best = std::move(plan);
    }
    if (best.G == 1) return CSMC_OK;
    const int G = best.G, ncol = best.ncol, K2 = best.K2;
    std::vector<int32_t> tab((size_t)G * SPLIT_HDR, 0);
    auto put = [&](const std::vector<int> &v) {
        const int32_t off = (int32_t)tab.size();
        tab.insert(tab.end(), v.begin(), v.end());
        return off;
    };
    for (int g = 0; g < G; ++g) {
        const SplitPart &sp = best.parts[g];
        const size_t h = (size_t)g * SPLIT_HDR;
        tab[h + 0] = sp.n_loc;
        tab[h + 1] = sp.n_own;
        tab[h + 2] = (int32_t)sp.slots.size();
        tab[h + 3] = (int32_t)sp.bnd.size();
        tab[h + 4] = (int32_t)sp.halo.size();
        tab[h + 5] = put(sp.col_start);
        tab[h + 6] = put(sp.own_cnt);
        tab[h + 7] = put(sp.phase_cnt);
        while (tab.size() % 4) tab.push_back(0);
        tab[h + 8] = (int32_t)tab.size();
        for (auto &q : sp.slots) tab.insert(tab.end(), q.begin(), q.end());
        std::vector<int> cs = sp.colslot;
        cs.insert(cs.end(), best.colbasis.begin(), best.colbasis.end());
        tab[h + 9] = put(cs);
        tab[h + 10] = put(sp.loc_dev);
        std::vector<int> ref(sp.n_loc);
        for (int i = 0; i < sp.n_loc; ++i) ref[i] = L.ref_of_dev[sp.loc_dev[i]];
        tab[h + 11] = put(ref);
        tab[h + 12] = put(sp.nbl);
        tab[h + 13] = put(sp.bnd);
        std::vector<int> hp;
        for (size_t j = 0; j < sp.halo.size(); ++j) {
            hp.push_back(sp.halo[j]);
            hp.push_back(sp.halo_src[j]);
        }
        while (tab.size() % 2) tab.push_back(0);
        tab[h + 14] = put(hp);
    }
    (void)ncol;
    (void)K2;
    const size_t nhx = 2 * (size_t)R * G * 3 * best.maxbnd, npx = 2 * (size_t)R * SPLIT_MAX * 2 * SPLIT_NV;
    HIPCHK(pt.split_tab.alloc(tab.size()));
    HIPCHK(pt.split_hx.alloc(nhx));
    HIPCHK(pt.split_px.alloc(npx));
    HIPCHK(pt.split_epoch.alloc(R));
    HIPCHK(hipMemcpyAsync(pt.split_tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, c->stream));
    // flags 0xFFFFFFFF: no epoch a run reaches
    HIPCHK(hipMemsetAsync(pt.split_hx.p, 0xFF, nhx * 8, c->stream));
    HIPCHK(hipMemsetAsync(pt.split_px.p, 0xFF, npx * 8, c->stream));
    HIPCHK(hipMemsetAsync(pt.split_epoch.p, 0, R * 4, c->stream));
    pt.split_G = G;
    pt.split_nor = nor;
    pt.split_spt = reg_spt(best.maxslots, b->k2c);
    pt.split_bmax = best.maxbnd;
    pt.split_lds = split_lds_bytes(best);
    return CSMC_OK;
}

WGArgs wg_base(csmc_ctx *c, Batch *b, uint64_t seed) {
    WGArgs a{};
    a.P = pools(c, b);
    a.spins = b->spins.p;
    a.k0 = (uint32_t)seed;
    a.k1 = (uint32_t)(seed >> 32);
    a.mcount = b->mcount.p;
    return a;
}

// fold finished event pairs into the running total (keeps the list short)
int drain_events(csmc_ctx *c, bool wait) {
    size_t k = 0;
    for (; k < c->events.size(); ++k) {
        auto &ev = c->events[k];
        if (wait) HIPCHK(hipEventSynchronize(ev.second));
        else if (hipEventQuery(ev.second) != hipSuccess) break;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev.first, ev.second));
        c->timed_ms += ms;
        ++c->timed_launches;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    c->events.erase(c->events.begin(), c->events.begin() + k);
    return CSMC_OK;
}

// Persistent launches in storage mode 0 with parts0 workgroups per system
// (sweeps, annealing, PT): the systems of a batch too large for LDS leave most
// CUs idle when each has one workgroup.  CSMC_PARTS0=1..8 forces the count
// (1: off); by default it is min(8, CUs / systems, sites per colour class /
// workgroup size), so every thread keeps a site per phase, when that is at
// least 2.
int setup_parts0(csmc_ctx *c, Batch *b) {
    PTHost &pt = b->pt;
    pt.parts0 = 1;
    pt.parts_ready = true;
    const int R = b->nsys;
    size_t bytes = 0;
    pt.parts_lm = choose_lm(c, b, &bytes);
    if (pt.parts_lm != 0) return CSMC_OK;  // spins in LDS: no parts
    int G = c->parts0;
    if (G == 0) {
        int per_colour = 0;  // the largest colour class of the batch's lattices
        for (int id : b->lat) {
            const HostLattice &L = *c->lats[id];
            for (int col = 0; col < L.ncol; ++col) {
                int n = 0;
                for (int r = L.colrun[col]; r < L.colrun[col + 1]; ++r) n += L.runs[r][1] - L.runs[r][0];
                per_colour = std::max(per_colour, n);
            }
        }
        const int threads = wg_threads_for(b->prec, b->self || b->gen, batch_sparse(c, b));
        G = std::min({8, c->ncu / R, per_colour / threads});
    }
    if (G < 2) return CSMC_OK;
    HIPCHK(pt.bar0.alloc(2 * (size_t)R));
    HIPCHK(pt.err0.alloc(1));
    HIPCHK(pt.psum0.alloc((size_t)R * 2 * G * 8));
    HIPCHK(pt.xcc0.alloc((size_t)R * G));
    HIPCHK(hipMemsetAsync(pt.bar0.p, 0, 2 * (size_t)R * 4, c->stream));
    HIPCHK(hipMemsetAsync(pt.err0.p, 0, 4, c->stream));
    pt.parts0 = G;
    return CSMC_OK;
}

// after a synchronised sweep or annealing call: whether a barrier wait of
// the parts gave up (the call's results are then not valid)
int parts_check(Batch *b) {
    PTHost &pt = b->pt;
    if (!pt.err0.p || (pt.parts0 < 2 && !(b->pack.use && b->pack.gmax > 1))) return CSMC_OK;
    uint32_t err = 0;
    HIPCHK(hipMemcpy(&err, pt.err0.p, 4, hipMemcpyDeviceToHost));
    if (err) {
        const uint32_t zero = 0;
        HIPCHK(hipMemcpy(pt.err0.p, &zero, 4, hipMemcpyHostToDevice));
        return fail(CSMC_EHIP, "a wait for another workgroup of the same system timed out");
    }
    return CSMC_OK;
}

// Packed schedule for sweeps and annealing in storage mode 0 (WGArgs::tasks).
// With one workgroup per system, a batch of lattices of uneven size takes as
// long as its costliest lattice while the CUs of the others idle.  A sweep of
// a system costs about ncol colour phases of ceil(largest colour class /
// (parts x workgroup size)) rounds of one site per thread (a phase has a
// fixed latency even when few threads have a site), plus one round per phase
// for the barrier when the system is split into parts.  The plan gives every
// CU one workgroup: for a target M it splits every system whose cost exceeds
// M into the fewest parts (at most 8) that bring it under M, on workgroups of
// one XCD (block b runs on XCD b % 8) that take nothing else first, and deals
// the other systems, costliest first, to the workgroup with the least cost
// so far (LPT scheduling); M rises from the total cost over the workgroups
// until a plan fits, and the plan is kept if its largest workgroup cost is
// below 0.9 of the costliest system's.  Each system's results do not depend
// on where, or in how many parts, it runs.
int pack_plan(csmc_ctx *c, Batch *b, int lm) {
    Batch::Pack &pk = b->pack;
    const int cap = std::min(c->lds_cap, b->lm_cap);
    if (pk.lm == lm && pk.cap == cap) return CSMC_OK;
    pk.lm = lm;
    pk.cap = cap;
    pk.use = false;
    const int R = b->nsys;
    const int threads = wg_threads_for(b->prec, b->self || b->gen, batch_sparse(c, b));
    const int nwg = c->ncu > 0 ? c->ncu : 256;
    if (!c->pack || lm != 0 || R < 2 || nwg % 8) return CSMC_OK;
    std::vector<int> ncol(R), maxcol(R, 0);
    for (int s = 0; s < R; ++s) {
        const HostLattice &L = *c->lats[b->lat[s]];
        ncol[s] = L.ncol;
        for (int col = 0; col < L.ncol; ++col) {
            int n = 0;
            for (int r = L.colrun[col]; r < L.colrun[col + 1]; ++r) n += L.runs[r][1] - L.runs[r][0];
            maxcol[s] = std::max(maxcol[s], n);
        }
    }
    auto cost = [&](int s, int g) {
        const int rounds = (maxcol[s] + g * threads - 1) / (g * threads);
        return (double)ncol[s] * (rounds + (g > 1 ? 0.25 : 0.0));
    };
    double total = 0, naive = 0;
    for (int s = 0; s < R; ++s) {
        total += cost(s, 1);
        naive = std::max(naive, cost(s, 1));
    }
    std::vector<int> order(R);
    for (int s = 0; s < R; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x, 1) > cost(y, 1); });
    std::vector<std::vector<int4>> best_list;
    double best_span = naive;
    int best_gmax = 1;
    for (double M = std::max(1.0, total / nwg); M < naive; M *= 1.05) {
        std::vector<int> G(R, 1);
        int used = 0, gmax = 1;
        bool ok = true;
        for (int s = 0; s < R && ok; ++s) {
            if (cost(s, 1) <= M) continue;
            int g = 2;
            while (g <= 8 && cost(s, g) > M) ++g;
            if (g > 8 || (maxcol[s] + threads - 1) / threads < g) ok = false;
            G[s] = g;
            used += g;
            gmax = std::max(gmax, g);
        }
        if (!ok || used > nwg) continue;
        std::vector<std::vector<int>> free_wg(8);
        for (int w = nwg - 1; w >= 0; --w) free_wg[w % 8].push_back(w);
        std::vector<std::vector<int4>> list(nwg);
        std::vector<double> load(nwg, 0.0);
        for (int s : order) {
            if (G[s] == 1) continue;
            int x = 0;
            for (int k = 1; k < 8; ++k)
                if (free_wg[k].size() > free_wg[x].size()) x = k;
            if ((int)free_wg[x].size() < G[s]) {
                ok = false;
                break;
            }
            for (int p = 0; p < G[s]; ++p) {
                const int w = free_wg[x].back();
                free_wg[x].pop_back();
                list[w].push_back(make_int4(s, p, G[s], 0));
                load[w] += cost(s, G[s]);
            }
        }
        if (!ok) continue;
        for (int s : order) {
            if (G[s] != 1) continue;
            int w = 0;
            for (int k = 1; k < nwg; ++k)
                if (load[k] < load[w]) w = k;
            list[w].push_back(make_int4(s, 0, 1, 0));
            load[w] += cost(s, 1);
        }
        double span = 0;
        for (double l : load) span = std::max(span, l);
        if (span < best_span) {
            best_span = span;
            best_list = std::move(list);
            best_gmax = gmax;
        }
        if (span <= M) break;  // the target is met: a larger one cannot do better
    }
    pk.makespan = best_span;
    pk.naive = naive;
    if (best_list.empty() || best_span > 0.9 * naive) return CSMC_OK;
    std::vector<int32_t> off(nwg + 1, 0);
    std::vector<int4> tasks;
    for (int w = 0; w < nwg; ++w) {
        off[w] = (int32_t)tasks.size();
        tasks.insert(tasks.end(), best_list[w].begin(), best_list[w].end());
    }
    off[nwg] = (int32_t)tasks.size();
    HIPCHK(pk.off.alloc(off.size()));
    HIPCHK(hipMemcpy(pk.off.p, off.data(), off.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(pk.tasks.alloc(tasks.size()));
    HIPCHK(hipMemcpy(pk.tasks.p, tasks.data(), tasks.size() * sizeof(int4), hipMemcpyHostToDevice));
    const int gmax = best_gmax;
    if (gmax > 1) {
        HIPCHK(pk.bar0.alloc(2 * (size_t)R));
        HIPCHK(pk.xcc0.alloc((size_t)R * gmax));
        HIPCHK(pk.psum0.alloc((size_t)R * 2 * gmax * 8));
        if (!b->pt.err0.p) {
            HIPCHK(b->pt.err0.alloc(1));
            HIPCHK(hipMemset(b->pt.err0.p, 0, 4));
        }
    }
    pk.nwg = nwg;
    pk.gmax = gmax;
    pk.use = true;
    if (c->verbose)
        fprintf(stderr, "csmc: packed schedule of %d systems on %d workgroups, largest cost %.0f rounds "
                        "per sweep (costliest system %.0f), up to %d parts\n", R, nwg, best_span, naive, gmax);
    return CSMC_OK;
}

int run_wg(csmc_ctx *c, Batch *b, const WGArgs &a) {
    size_t lds_bytes = 0;
    int lm = choose_lm(c, b, &lds_bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, c->stream));
    }
    bool split_done = false;
    if (lm == 3 && a.mode == WG_PT_CHAIN && b->pt.split_G > 1) {
        PTHost &pt = b->pt;
        bool diag = true;
        for (int id : b->lat) diag = diag && (c->h_lat[id].flags & LAT_DIAG);
        WGArgs s = a;
        s.split.G = pt.split_G;
        s.split.nor = pt.split_nor;
        s.split.R = b->nsys;
        s.split.bmax = pt.split_bmax;
        s.split.tab = pt.split_tab.p;
        s.split.hx = pt.split_hx.p;
        s.split.px = pt.split_px.p;
        s.split.epoch = pt.split_epoch.p;
        hipError_t e = launch_wg_split(b->k2c, pt.split_spt, diag, pt.split_G, b->nsys, pt.split_lds, s,
                                       c->stream);
        if (e == hipSuccess) {
            split_done = true;
        } else {
            (void)hipGetLastError();  // not co-resident: the whole-system kernel from now on
            pt.split_G = 1;
        }
    }
    if (split_done) {
    } else if (lm == 3) {
        bool diag = true, two = true;
        const int spt = reg_spt(b->nslots, b->k2c);
        for (int id : b->lat) {
            diag = diag && (c->h_lat[id].flags & LAT_DIAG);
            // two full colour classes of spt / 2 slots each
            two = two && (c->h_lat[id].flags & LAT_TWO_COLOUR) && c->h_lat[id].nslots == spt;
        }
        HIPCHK(launch_wg_reg(b->k2c, spt, diag, two, b->nsys, lds_bytes, a, c->stream));
    } else {
        PTHost &pt = b->pt;
        bool parts_done = false;
        const bool sparse = batch_sparse(c, b);
        // chosen when the batch first runs in storage mode 0, also when an
        // earlier choice was made for an LDS mode (the autotuner's caps)
        if (lm == 0 && (!pt.parts_ready || pt.parts_lm != 0)) {
            const int rc = setup_parts0(c, b);
            if (rc) return rc;
        }
        if (lm == 0 && (a.mode == WG_SWEEPS || a.mode == WG_ANNEAL)) {
            const int rc = pack_plan(c, b, lm);
            if (rc) return rc;
        }
        if (lm == 0 && (a.mode == WG_SWEEPS || a.mode == WG_ANNEAL) && b->pack.use) {
            Batch::Pack &pk = b->pack;
            WGArgs s = a;
            s.tasks = pk.tasks.p;
            s.task_off = pk.off.p;
            s.parts0 = pk.gmax;
            s.bar0 = pk.bar0.p;
            s.psum0 = pk.psum0.p;
            s.xcc0 = pk.xcc0.p;
            s.err0 = pt.err0.p;
            s.spin_limit = c->spin_limit;
            if (pk.gmax > 1) HIPCHK(hipMemsetAsync(pk.bar0.p, 0, pk.bar0.n * sizeof(uint32_t), c->stream));
            hipError_t e = launch_wg(b->prec, lm, b->self, b->gen, c->specialise ? b->k2c : 0, sparse, pk.nwg,
                                     lds_bytes, s, c->stream);
            if (e == hipSuccess) {
                parts_done = true;
            } else {
                (void)hipGetLastError();  // not co-resident: the plain schedule from now on
                pk.use = false;
            }
        }
        if (!parts_done && lm == 0 && pt.parts0 > 1) {
            // a launch that gave up on a wait may have left arrivals behind
            HIPCHK(hipMemsetAsync(pt.bar0.p, 0, pt.bar0.n * sizeof(uint32_t), c->stream));
            WGArgs s = a;
            s.parts0 = pt.parts0;
            s.bar0 = pt.bar0.p;
            s.psum0 = pt.psum0.p;
            s.xcc0 = pt.xcc0.p;
            // PT waits report through the PT state's flag, checked after
            // every csmc_pt_advance; sweeps and annealing through err0
            s.err0 = (a.mode == WG_PT || a.mode == WG_PT_CHAIN) ? pt.err.p : pt.err0.p;
            s.spin_limit = c->spin_limit;  // only the PT paths set it
            const size_t tab = part_table_bytes(c, b, pt.parts0);
            s.ptab = tab > 0;
            hipError_t e = launch_wg(b->prec, lm, b->self, b->gen, c->specialise ? b->k2c : 0, sparse,
                                     b->nsys * pt.parts0, tab > 0 ? tab : lds_bytes, s, c->stream);
            if (e == hipSuccess) {
                parts_done = true;
            } else {
                (void)hipGetLastError();  // not co-resident: one workgroup per system from now on
                pt.parts0 = 1;
            }
        }
        if (!parts_done) {
            const size_t tab = lm == 0 ? part_table_bytes(c, b, 1) : 0;
            WGArgs s = a;
            s.ptab = tab > 0;
            HIPCHK(launch_wg(b->prec, lm, b->self, b->gen, c->specialise ? b->k2c : 0, sparse, b->nsys,
                             tab > 0 ? tab : lds_bytes, s, c->stream));
        }
    }
    if (c->timing) {
        HIPCHK(hipEventRecord(e1, c->stream));
        c->events.push_back({e0, e1});
        if (c->events.size() > 256) return drain_events(c, false);
    }
    return CSMC_OK;
}

}  // namespace

namespace {
int sweep_passes(csmc_ctx *c, Batch *b, int32_t kind, const double *T, const double *sigma, int32_t n_sweeps,
                 uint64_t seed, double *accepted);
bool tiles_ready(csmc_ctx *c, Batch *b);
int constraint_run(csmc_ctx *c, Batch *b, uint64_t seed, int64_t n_sweeps);
}  // namespace

extern "C" {

int32_t csmc_version(void) { return 1; }

const char *csmc_last_error(void) { return g_err.c_str(); }

int32_t csmc_device_count(int32_t *n) {
    int v = 0;
    hipError_t e = hipGetDeviceCount(&v);
    if (e != hipSuccess) v = 0;
    if (n) *n = v;
    return CSMC_OK;
}

int32_t csmc_ctx_create(int32_t device, csmc_ctx **out) {
    if (!out) return fail(CSMC_EINVAL, "null out pointer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(CSMC_EHIP, "no HIP device available: the engine has no CPU fallback");
    if (device < 0 || device >= n) return fail(CSMC_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    if (int rc = scratch_setup(device)) return rc;
    auto *c = new csmc_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(CSMC_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->max_lds = max_lds_bytes();
    const char *lc = getenv("CSMC_LDS");
    if (lc && lc[0] >= '0' && lc[0] <= '3') c->lds_cap = lc[0] - '0';
    const char *sp = getenv("CSMC_SPECIALISE");
    if (sp && sp[0] == '0') c->specialise = false;
    const char *ch = getenv("CSMC_PT_CHAIN");
    if (ch && ch[0] == '0') c->pt_chain = false;
    // a smaller limit makes the waits give up early: for tests of the
    // failure path (every wave still ends, the call reports the error)
    const char *lim = getenv("CSMC_SPIN_LIMIT");
    if (lim && lim[0] >= '0' && lim[0] <= '9') c->spin_limit = (uint32_t)strtoul(lim, nullptr, 10);
    const char *sl = getenv("CSMC_SPLIT");
    if (sl && std::string(sl) == "auto") c->split = 0;
    else if (sl && sl[0] >= '0' && sl[0] <= '8' && sl[1] == 0) c->split = sl[0] == '0' ? 1 : sl[0] - '0';
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess)
            c->ncu = ncu;
    }
    const char *sf = getenv("CSMC_SSF");
    if (sf && std::string(sf) == "blas") c->ssf_blas = true;
    const char *p0 = getenv("CSMC_PARTS0");
    if (p0 && p0[0] >= '1' && p0[0] <= '8' && p0[1] == 0) c->parts0 = p0[0] - '0';
    const char *pk = getenv("CSMC_PACK");
    if (pk && pk[0] == '0') c->pack = false;
    const char *ptb = getenv("CSMC_PARTTAB");
    if (ptb && ptb[0] == '0') c->parttab = false;
    const char *tl = getenv("CSMC_TILE");
    if (tl && tl[0] == '0') c->tile = false;
    const char *tb = getenv("CSMC_TILE_LDS");
    if (tb && tb[0] >= '1' && tb[0] <= '9') c->tile_lds = c->tile_lds1 = (size_t)strtoul(tb, nullptr, 10) << 10;
    const char *tb1 = getenv("CSMC_TILE_LDS1");
    if (tb1 && tb1[0] >= '1' && tb1[0] <= '9') c->tile_lds1 = (size_t)strtoul(tb1, nullptr, 10) << 10;
    const char *tsw = getenv("CSMC_TILE_SWEEPS");
    if (tsw && tsw[0] >= '1' && tsw[0] <= '8' && tsw[1] == 0) c->tile_sweeps = tsw[0] - '0';
    const char *dt = getenv("CSMC_DENSE_TENSORS");
    if (dt && dt[0] == '1') c->dense_tensors = true;
    const char *vb = getenv("CSMC_VERBOSE");
    c->verbose = vb && vb[0] == '1';
    if (c->verbose) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess)
            fprintf(stderr, "csmc: device %d %s (%s), %d CUs, max LDS per block %d B%s\n", device,
                    prop.name, prop.gcnArchName, prop.multiProcessorCount, c->max_lds,
                    c->lds_cap < 3 ? " (storage modes capped)" : "");
    }
    *out = c;
    return CSMC_OK;
}

int32_t csmc_ctx_destroy(csmc_ctx *c) {
    CHECK_CTX(c);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->batches.clear();
    blas_destroy(c->blas);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return CSMC_OK;
}

int32_t csmc_ctx_sync(csmc_ctx *c) {
    CHECK_CTX(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_timing(csmc_ctx *c, int32_t enable) {
    CHECK_CTX(c);
    int rc = drain_events(c, true);
    if (rc) return rc;
    c->timing = enable != 0;
    c->timed_ms = 0.0;
    c->timed_launches = 0;
    return CSMC_OK;
}

int32_t csmc_timing_read(csmc_ctx *c, double *total_ms, int64_t *launches) {
    CHECK_CTX(c);
    int rc = drain_events(c, true);
    if (rc) return rc;
    if (total_ms) *total_ms = c->timed_ms;
    if (launches) *launches = c->timed_launches;
    return CSMC_OK;
}

static void fill_info(const HostLattice &L, csmc_lattice_info *info) {
    info->n_sites = L.N;
    info->dim = L.dim;
    info->n_basis = L.nb;
    info->n2 = L.n2;
    info->n3 = L.n3;
    info->n4 = L.n4;
    info->k2 = L.K2;
    info->k3 = L.K3;
    info->k4 = L.K4;
    info->n_colours = L.ncol;
    info->has_self = L.has_self;
    info->S = L.S;
}

int32_t csmc_build_tables(const csmc_unit_cell *uc, const int32_t *shape, int32_t periodic,
                          double S, csmc_lattice_info *info, double *pos, double *field,
                          double *onsite, int64_t *bsite, double *bmat, int64_t *csite,
                          double *ctens, int64_t *qsite, double *qtens, int32_t *colour) {
    HostLattice L;
    std::string err;
    int rc = build_lattice(uc, shape, periodic, S, L, err);
    if (rc) return fail(rc, err);
    if (info) fill_info(L, info);
    export_tables(L, pos, field, onsite, bsite, bmat, csite, ctens, qsite, qtens);
    if (colour)
        for (long i = 0; i < L.N; ++i) colour[i] = L.colour[i];
    return CSMC_OK;
}

int32_t csmc_split_plan(const csmc_unit_cell *uc, const int32_t *shape, int32_t periodic,
                        int32_t parts, int32_t or_sweeps, int64_t *stats) {
    if (!stats || parts < 1 || parts > SPLIT_MAX || or_sweeps < 0)
        return fail(CSMC_EINVAL, "split plan: parts in 1..8, or_sweeps >= 0");
    HostLattice L;
    std::string err;
    int rc = build_lattice(uc, shape, periodic, 0.5, L, err);
    if (rc) return fail(rc, err);
    SplitPlan plan;
    if (!build_split(L, parts, or_sweeps, REG_THREADS, plan))
        return fail(CSMC_EINVAL, "the lattice cannot be split into these parts");
    stats[0] = plan.G;
    stats[1] = plan.P;
    stats[2] = plan.maxloc;
    stats[3] = plan.maxslots;
    stats[4] = plan.work_max;
    stats[5] = plan.work_full;
    stats[6] = check_split(L, plan);
    for (int g = 0; g < plan.G; ++g) {
        const SplitPart &sp = plan.parts[g];
        stats[7 + 4 * g] = sp.n_loc;
        stats[8 + 4 * g] = sp.n_own;
        stats[9 + 4 * g] = (int64_t)sp.halo.size();
        stats[10 + 4 * g] = (int64_t)sp.bnd.size();
    }
    return CSMC_OK;
}

int32_t csmc_tile_plan(const csmc_unit_cell *uc, const int32_t *shape, int32_t periodic, int32_t prec,
                       int64_t lds_budget, int32_t sweeps, int32_t *head, int32_t *tab, int32_t tab_cap,
                       int32_t *tab_len) {
    if (!head || !tab_len || (prec != 32 && prec != 64) || lds_budget < 0 || sweeps < 1 || sweeps > 8)
        return fail(CSMC_EINVAL, "tile plan: head and tab_len needed, prec 32 or 64, sweeps 1..8");
    HostLattice L;
    std::string err;
    int rc = build_lattice(uc, shape, periodic, 0.5, L, err);
    if (rc) return fail(rc, err);
    const TilePlan P = plan_tiles(L, 3 * (prec == 32 ? 4 : 8), (size_t)lds_budget, sweeps);
    static_assert(sizeof(TileDev) == CSMC_TILE_HEAD * 4, "tile plan header");
    if (!P.ok) {
        std::memset(head, 0, sizeof(TileDev));
        *tab_len = 0;
        g_err = P.why;
        return CSMC_OK;
    }
    std::memcpy(head, &P.d, sizeof(TileDev));
    *tab_len = (int32_t)P.tab.size();
    if (tab && tab_cap >= (int32_t)P.tab.size()) std::memcpy(tab, P.tab.data(), P.tab.size() * 4);
    return CSMC_OK;
}

int32_t csmc_lattice_create(csmc_ctx *c, const csmc_unit_cell *uc, const int32_t *shape,
                            int32_t periodic, double S, int32_t *lattice_id) {
    CHECK_CTX(c);
    auto L = std::make_unique<HostLattice>();
    std::string err;
    int rc = build_lattice(uc, shape, periodic, S, *L, err);
    if (rc) return fail(rc, err);
    append_lattice(c, *L);
    c->lats.push_back(std::move(L));
    if (lattice_id) *lattice_id = (int32_t)c->lats.size() - 1;
    return CSMC_OK;
}

int32_t csmc_lattice_info_get(csmc_ctx *c, int32_t id, csmc_lattice_info *info) {
    CHECK_CTX(c);
    if (id < 0 || id >= (int)c->lats.size()) return fail(CSMC_ESTATE, "unknown lattice id");
    fill_info(*c->lats[id], info);
    return CSMC_OK;
}

int32_t csmc_lattice_tables(csmc_ctx *c, int32_t id, double *pos, double *field, double *onsite,
                            int64_t *bsite, double *bmat, int64_t *csite, double *ctens,
                            int64_t *qsite, double *qtens) {
    CHECK_CTX(c);
    if (id < 0 || id >= (int)c->lats.size()) return fail(CSMC_ESTATE, "unknown lattice id");
    export_tables(*c->lats[id], pos, field, onsite, bsite, bmat, csite, ctens, qsite, qtens);
    return CSMC_OK;
}

int32_t csmc_lattice_colouring(csmc_ctx *c, int32_t id, int32_t *colour) {
    CHECK_CTX(c);
    if (id < 0 || id >= (int)c->lats.size()) return fail(CSMC_ESTATE, "unknown lattice id");
    const HostLattice &L = *c->lats[id];
    for (long i = 0; i < L.N; ++i) colour[i] = L.colour[i];
    return CSMC_OK;
}

int32_t csmc_lattice_update_couplings(csmc_ctx *c, int32_t id, const csmc_unit_cell *uc) {
    CHECK_CTX(c);
    if (id < 0 || id >= (int)c->lats.size()) return fail(CSMC_ESTATE, "unknown lattice id");
    HostLattice &L = *c->lats[id];
    std::string err;
    int rc = update_couplings(uc, L, err);
    if (rc) return fail(rc, err);
    std::vector<double> dc = device_coef(L);
    std::copy(dc.begin(), dc.end(), c->h_coef.begin() + c->h_lat[id].coef_off);
    c->h_lat[id].flags = coupling_flags(L) | (gradient_is_local_field(L) ? LAT_GRAD_LF : 0) |
                         (c->h_lat[id].flags & (LAT_COLBASIS | LAT_TWO_COLOUR | LAT_SELF_Q4));
    const std::vector<int32_t> pat = tensor_patterns(L, c->dense_tensors);
    for (int32_t v : pat)
        if (v == TP_DENSE) c->h_lat[id].flags |= LAT_DENSE_T;
    std::copy(pat.begin(), pat.end(), c->h_i32.begin() + c->h_lat[id].pat_off);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->dirty = true;
    return CSMC_OK;
}

int32_t csmc_batch_create(csmc_ctx *c, int32_t n, const int32_t *lat_ids, int32_t prec,
                          int32_t *batch_id) {
    CHECK_CTX(c);
    if (n < 1 || !lat_ids) return fail(CSMC_EINVAL, "batch needs at least one system");
    if (prec != 32 && prec != 64) return fail(CSMC_EINVAL, "precision must be 32 or 64");
    auto b = std::make_unique<Batch>();
    b->prec = prec;
    b->nsys = n;
    int64_t off = 0, sites = 0;
    for (int s = 0; s < n; ++s) {
        int id = lat_ids[s];
        if (id < 0 || id >= (int)c->lats.size())
            return fail(CSMC_ESTATE, "unknown lattice id in batch");
        const HostLattice &L = *c->lats[id];
        SysDev sd;
        sd.lat = id;
        sd.stream = s;
        sd.spin_off = off;
        sd.site_off = sites;
        b->sys.push_back(sd);
        b->lat.push_back(id);
        off += L.N + 1;  // one zero pad site per system
        sites += L.N;
        b->maxN = std::max<int>(b->maxN, (int)L.N);
        b->maxK2 = std::max<int>(b->maxK2, L.K2);
        b->k2c = (s == 0 || b->k2c == L.K2) ? L.K2 : 0;
        b->nslots = std::max(b->nslots, c->h_lat[id].nslots);
        b->self |= (L.has_self != 0);
        b->gen |= (L.K3 > 0 || L.K4 > 0);
    }
    b->total = sites;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(b->sys_d.alloc(n));
    HIPCHK(hipMemcpy(b->sys_d.p, b->sys.data(), n * sizeof(SysDev), hipMemcpyHostToDevice));
    HIPCHK(b->spins.alloc((size_t)off * b->spin_bytes()));
    HIPCHK(hipMemset(b->spins.p, 0, (size_t)off * b->spin_bytes()));
    HIPCHK(b->mcount.alloc(n));
    HIPCHK(hipMemset(b->mcount.p, 0, n * sizeof(uint32_t)));
    HIPCHK(b->T.alloc(n));
    HIPCHK(b->sigma.alloc(n));
    HIPCHK(b->acc.alloc(n));
    HIPCHK(b->accU.alloc(n));
    HIPCHK(b->stage.alloc(std::max<size_t>((size_t)off * 3, (size_t)8 * n)));
    int id = c->next_batch++;
    c->batches[id] = std::move(b);
    if (batch_id) *batch_id = id;
    return CSMC_OK;
}

int32_t csmc_batch_destroy(csmc_ctx *c, int32_t id) {
    CHECK_CTX(c);
    (void)hipStreamSynchronize(c->stream);
    if (!c->batches.erase(id)) return fail(CSMC_ESTATE, "unknown batch id");
    return CSMC_OK;
}

int32_t csmc_set_spins(csmc_ctx *c, int32_t id, int32_t sys, const double *spins) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (sys < 0 || sys >= b->nsys || !spins) return fail(CSMC_EINVAL, "bad system index or pointer");
    int rc = ensure_pools(c);
    if (rc) return rc;
    int N = (int)c->lats[b->lat[sys]]->N;
    double *st = b->stage.p + 3 * b->sys[sys].spin_off;
    HIPCHK(hipMemcpyAsync(st, spins, (size_t)N * 3 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_scatter_spins(b->prec, pools(c, b), b->spins.p, sys, N, st, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_get_spins(csmc_ctx *c, int32_t id, int32_t sys, double *spins) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (sys < 0 || sys >= b->nsys || !spins) return fail(CSMC_EINVAL, "bad system index or pointer");
    int rc = ensure_pools(c);
    if (rc) return rc;
    int N = (int)c->lats[b->lat[sys]]->N;
    double *st = b->stage.p + 3 * b->sys[sys].spin_off;
    HIPCHK(launch_gather_spins(b->prec, pools(c, b), b->spins.p, sys, N, st, c->stream));
    HIPCHK(hipMemcpyAsync(spins, st, (size_t)N * 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_randomize_spins(csmc_ctx *c, int32_t id, uint64_t seed) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    int rc = ensure_pools(c);
    if (rc) return rc;
    HIPCHK(launch_randomize(b->prec, pools(c, b), b->spins.p, b->nsys, b->maxN, (uint32_t)seed,
                            (uint32_t)(seed >> 32), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_local_field(csmc_ctx *c, int32_t id, int32_t sys, double *H) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (sys < 0 || sys >= b->nsys || !H) return fail(CSMC_EINVAL, "bad system index or pointer");
    int rc = ensure_pools(c);
    if (rc) return rc;
    int N = (int)c->lats[b->lat[sys]]->N;
    double *st = b->stage.p + 3 * b->sys[sys].spin_off;
    HIPCHK(launch_site_eval(b->prec, pools(c, b), b->spins.p, sys, N, st, nullptr, c->stream));
    HIPCHK(hipMemcpyAsync(H, st, (size_t)N * 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_site_energy(csmc_ctx *c, int32_t id, int32_t sys, double *e) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (sys < 0 || sys >= b->nsys || !e) return fail(CSMC_EINVAL, "bad system index or pointer");
    int rc = ensure_pools(c);
    if (rc) return rc;
    int N = (int)c->lats[b->lat[sys]]->N;
    double *st = b->stage.p + 3 * b->sys[sys].spin_off;
    HIPCHK(launch_site_eval(b->prec, pools(c, b), b->spins.p, sys, N, nullptr, st, c->stream));
    HIPCHK(hipMemcpyAsync(e, st, (size_t)N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_site_eval(csmc_ctx *c, int32_t id, int32_t sys, int32_t n_set, const int64_t *set_sites,
                       const double *set_spins, int32_t n_at, const int64_t *at_sites, double *H, double *e) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (sys < 0 || sys >= b->nsys) return fail(CSMC_EINVAL, "bad system index");
    if (n_set < 0 || n_at < 0 || n_set > CSMC_SITE_EVAL_MAX || n_at > CSMC_SITE_EVAL_MAX)
        return fail(CSMC_EINVAL, "site counts must be in 0..CSMC_SITE_EVAL_MAX");
    if ((n_set && (!set_sites || !set_spins)) || (n_at && (!at_sites || (!H && !e))))
        return fail(CSMC_EINVAL, "null pointer");
    int rc = ensure_pools(c);
    if (rc) return rc;
    const HostLattice &L = *c->lats[b->lat[sys]];
    // device indices (site order on the device is (colour, basis, reference))
    std::vector<int32_t> idx((size_t)n_set + n_at);
    for (int k = 0; k < n_set + n_at; ++k) {
        const int64_t r = k < n_set ? set_sites[k] : at_sites[k - n_set];
        if (r < 0 || r >= L.N) return fail(CSMC_EBOUNDS, "site index out of range");
        idx[k] = L.dev_of_ref[r];
    }
    // one staging block: indices, then the values, then the outputs
    const size_t bi = 4 * idx.size(), bv = 24 * (size_t)n_set, bo = 32 * (size_t)n_at;
    const size_t ov = (bi + 7) & ~(size_t)7, oo = ov + bv;
    if (b->sites.n < oo + bo) HIPCHK(b->sites.alloc(oo + bo));
    unsigned char *d = b->sites.p;
    HIPCHK(hipMemcpyAsync(d, idx.data(), bi, hipMemcpyHostToDevice, c->stream));
    if (n_set) HIPCHK(hipMemcpyAsync(d + ov, set_spins, bv, hipMemcpyHostToDevice, c->stream));
    double *dH = reinterpret_cast<double *>(d + oo), *de = dH + 3 * (size_t)n_at;
    HIPCHK(launch_site_eval_at(b->prec, pools(c, b), b->spins.p, sys, reinterpret_cast<const int32_t *>(d),
                               reinterpret_cast<const double *>(d + ov), n_set,
                               reinterpret_cast<const int32_t *>(d) + n_set, n_at, H ? dH : nullptr,
                               e ? de : nullptr, c->stream));
    if (n_at && H) HIPCHK(hipMemcpyAsync(H, dH, 24 * (size_t)n_at, hipMemcpyDeviceToHost, c->stream));
    if (n_at && e) HIPCHK(hipMemcpyAsync(e, de, 8 * (size_t)n_at, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_total_energy(csmc_ctx *c, int32_t id, double *E, double *parts) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!E) return fail(CSMC_EINVAL, "null output");
    int rc = ensure_pools(c);
    if (rc) return rc;
    double *dE = b->stage.p, *dP = b->stage.p + b->nsys;
    HIPCHK(launch_energy(b->prec, pools(c, b), b->spins.p, b->nsys, dE, dP, c->stream));
    HIPCHK(hipMemcpyAsync(E, dE, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    if (parts) HIPCHK(hipMemcpyAsync(parts, dP, b->nsys * 5 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_magnetization(csmc_ctx *c, int32_t id, double *M) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!M) return fail(CSMC_EINVAL, "null output");
    int rc = ensure_pools(c);
    if (rc) return rc;
    HIPCHK(launch_magnet(b->prec, pools(c, b), b->spins.p, b->nsys, b->stage.p, c->stream));
    HIPCHK(hipMemcpyAsync(M, b->stage.p, b->nsys * 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

static bool valid_kind(int k) { return k >= CSMC_OVERRELAX && k <= CSMC_DETERMINISTIC; }

int32_t csmc_sweep(csmc_ctx *c, int32_t id, int32_t kind, const double *T, double *sigma,
                   int32_t n_sweeps, uint64_t seed, double *accepted) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (kind == CSMC_CONSTRAINT_ADAPTIVE) {
        if (!b->cons) return fail(CSMC_ESTATE, "no constraint on this batch (csmc_constraint_set)");
        if (!T) return fail(CSMC_EINVAL, "Metropolis sweeps need temperatures");
        if (n_sweeps < 0) return fail(CSMC_EINVAL, "n_sweeps < 0");
        int rc = ensure_pools(c);
        if (rc) return rc;
        std::vector<double> sg(b->nsys, b->cons_sigma0);
        if (sigma) sg.assign(sigma, sigma + b->nsys);
        HIPCHK(hipMemcpyAsync(b->T.p, T, b->nsys * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b->sigma.p, sg.data(), b->nsys * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(b->cons_total.p, 0, b->nsys * 8, c->stream));
        rc = constraint_run(c, b, seed, n_sweeps);
        if (rc) return rc;
        if (accepted) HIPCHK(hipMemcpyAsync(accepted, b->cons_total.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
        if (sigma) HIPCHK(hipMemcpyAsync(sigma, b->sigma.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return CSMC_OK;
    }
    if (!valid_kind(kind)) return fail(CSMC_EINVAL, "unknown sweep kind");
    if (b->passes && kind != CSMC_CONE_ADAPTIVE)  // chosen by csmc_autotune
        return sweep_passes(c, b, kind, T, sigma, n_sweeps, seed, accepted);
    if (n_sweeps < 0) return fail(CSMC_EINVAL, "n_sweeps < 0");
    bool metro = kind != CSMC_OVERRELAX && kind != CSMC_DETERMINISTIC;
    if (metro && !T) return fail(CSMC_EINVAL, "Metropolis sweeps need temperatures");
    int rc = ensure_pools(c);
    if (rc) return rc;
    if (T) HIPCHK(hipMemcpyAsync(b->T.p, T, b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(hipMemsetAsync(b->T.p, 0, b->nsys * 8, c->stream));
    if (sigma) HIPCHK(hipMemcpyAsync(b->sigma.p, sigma, b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    WGArgs a = wg_base(c, b, seed);
    a.mode = WG_SWEEPS;
    a.kind = kind;
    a.n_sweeps = n_sweeps;
    a.T_sys = b->T.p;
    a.sigma_sys = sigma ? b->sigma.p : nullptr;
    a.acc_sys = b->acc.p;
    rc = run_wg(c, b, a);
    if (rc) return rc;
    if (accepted) HIPCHK(hipMemcpyAsync(accepted, b->acc.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    if (sigma) HIPCHK(hipMemcpyAsync(sigma, b->sigma.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return parts_check(b);
}

int32_t csmc_sweep_plain(csmc_ctx *c, int32_t id, int32_t kind, const double *T, double *sigma,
                         int32_t n_sweeps, uint64_t seed, double *accepted) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!valid_kind(kind)) return fail(CSMC_EINVAL, "unknown sweep kind");
    bool metro = kind != CSMC_OVERRELAX && kind != CSMC_DETERMINISTIC;
    if (metro && !T) return fail(CSMC_EINVAL, "Metropolis sweeps need temperatures");
    int rc = ensure_pools(c);
    if (rc) return rc;
    std::vector<uint32_t> ctr(b->nsys);
    HIPCHK(hipMemcpyAsync(ctr.data(), b->mcount.p, b->nsys * 4, hipMemcpyDeviceToHost, c->stream));
    if (T) HIPCHK(hipMemcpyAsync(b->T.p, T, b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<double> sg(b->nsys, 60.0);
    if (sigma) sg.assign(sigma, sigma + b->nsys);
    HIPCHK(hipMemcpyAsync(b->sigma.p, sg.data(), b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(b->accU.p, 0, b->nsys * 4, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<double> acc(b->nsys, 0.0);
    PassArgs a{};
    a.P = pools(c, b);
    a.spins = b->spins.p;
    a.kind = kind;
    a.k0 = (uint32_t)seed;
    a.k1 = (uint32_t)(seed >> 32);
    a.ctr = nullptr;
    a.T_sys = b->T.p;
    a.sigma_sys = b->sigma.p;
    a.acc = b->accU.p;
    for (int s = 0; s < b->nsys; ++s) {
        const HostLattice &L = *c->lats[b->lat[s]];
        a.sys0 = s;
        for (int sw = 0; sw < n_sweeps; ++sw) {
            a.ctr0 = ctr[s];
            for (int col = 0; col < L.ncol; ++col) {
                int n_col = L.runs[L.colrun[col + 1] - 1][1] - L.runs[L.colrun[col]][0];
                a.colour = col;
                HIPCHK(launch_pass(b->prec, b->self, b->gen, 1, n_col, a, c->stream));
            }
            if (metro) {
                ++ctr[s];
                if (kind == CSMC_CONE_ADAPTIVE) {
                    // metropolis_adaptive! width update after the sweep
                    uint32_t got = 0;
                    HIPCHK(hipMemcpyAsync(&got, b->accU.p + s, 4, hipMemcpyDeviceToHost, c->stream));
                    HIPCHK(hipMemsetAsync(b->accU.p + s, 0, 4, c->stream));
                    HIPCHK(hipStreamSynchronize(c->stream));
                    acc[s] += got;
                    double av = (double)got / (double)L.N;
                    double f = 0.5 / std::max(1.0 - av, 0.05);
                    sg[s] = std::min(std::max(sg[s] * f, 0.0), 100.0);
                    HIPCHK(hipMemcpyAsync(b->sigma.p + s, &sg[s], 8, hipMemcpyHostToDevice, c->stream));
                    HIPCHK(hipStreamSynchronize(c->stream));
                }
            }
        }
    }
    std::vector<uint32_t> got(b->nsys);
    HIPCHK(hipMemcpyAsync(got.data(), b->accU.p, b->nsys * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(b->mcount.p, ctr.data(), b->nsys * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int s = 0; s < b->nsys; ++s) acc[s] += got[s];
    if (accepted) std::copy(acc.begin(), acc.end(), accepted);
    if (sigma) std::copy(sg.begin(), sg.end(), sigma);
    return CSMC_OK;
}

int32_t csmc_sweep_passes(csmc_ctx *c, int32_t id, int32_t kind, const double *T, const double *sigma,
                          int32_t n_sweeps, uint64_t seed, double *accepted) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    return sweep_passes(c, b, kind, T, sigma, n_sweeps, seed, accepted);
}

int32_t csmc_batch_tiles(csmc_ctx *c, int32_t id, int32_t *tiles, int32_t *sweeps) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!tiles) return fail(CSMC_EINVAL, "null output");
    int rc = ensure_pools(c);
    if (rc) return rc;
    *tiles = (c->tile && tiles_ready(c, b)) ? b->tiles.one.max_tiles : 0;
    if (sweeps) *sweeps = *tiles ? b->tiles.sweeps : 1;
    if (!*tiles) g_err = c->tile ? b->tiles.why : "tiles switched off (CSMC_TILE=0)";
    return CSMC_OK;
}

int32_t csmc_autotune(csmc_ctx *c, int32_t id, int32_t kind, const double *T, int32_t n_sweeps, double *times,
                      int32_t *choice) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!valid_kind(kind)) return fail(CSMC_EINVAL, "unknown sweep kind");
    if (n_sweeps < 1) return fail(CSMC_EINVAL, "n_sweeps < 1");
    const bool metro = kind != CSMC_OVERRELAX && kind != CSMC_DETERMINISTIC;
    if (metro && !T) return fail(CSMC_EINVAL, "Metropolis sweeps need temperatures");
    int rc = ensure_pools(c);
    if (rc) return rc;
    // the state the trials start from, restored after each
    DevBuf<unsigned char> spins0;
    DevBuf<uint32_t> ctr0;
    HIPCHK(spins0.alloc(b->spins.n));
    HIPCHK(ctr0.alloc(b->mcount.n));
    HIPCHK(hipMemcpyAsync(spins0.p, b->spins.p, b->spins.n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(ctr0.p, b->mcount.p, b->mcount.n * 4, hipMemcpyDeviceToDevice, c->stream));
    auto restore = [&]() -> int {
        HIPCHK(hipMemcpyAsync(b->spins.p, spins0.p, b->spins.n, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b->mcount.p, ctr0.p, b->mcount.n * 4, hipMemcpyDeviceToDevice, c->stream));
        return CSMC_OK;
    };
    const int cap0 = b->lm_cap;
    const bool passes0 = b->passes;
    const bool timing0 = c->timing;
    c->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // candidates 0..3: the persistent kernel capped at storage mode 3..0
    // (a cap that selects a mode already tried is skipped); 4: batch-wide
    // colour passes (not for the adaptive cone)
    double best = 1e300;
    int best_k = -1, seen_lm[4] = {-1, -1, -1, -1};
    for (int k = 0; k < 5; ++k) {
        if (times) times[k] = std::nan("");
        if (k < 4) {
            b->lm_cap = 3 - k;
            b->passes = false;
            size_t bytes = 0;
            const int lm = choose_lm(c, b, &bytes);
            bool dup = false;
            for (int j = 0; j < k; ++j) dup = dup || seen_lm[j] == lm;
            seen_lm[k] = lm;
            if (dup) continue;
        } else {
            if (kind == CSMC_CONE_ADAPTIVE) continue;
            b->lm_cap = 3;
            b->passes = true;
        }
        double ms_best = 1e300;
        for (int rep = 0; rep < 2 && rc == CSMC_OK; ++rep) {  // the first run pays one-time costs
            rc = restore();
            if (rc) break;
            HIPCHK(hipEventRecord(e0, c->stream));
            rc = b->passes ? sweep_passes(c, b, kind, T, nullptr, n_sweeps, 1, nullptr)
                           : csmc_sweep(c, id, kind, T, nullptr, n_sweeps, 1, nullptr);
            if (rc) break;
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            ms_best = std::min(ms_best, (double)ms);
        }
        if (rc) break;
        if (times) times[k] = ms_best;
        if (ms_best < best) {
            best = ms_best;
            best_k = k;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->timing = timing0;
    const int rc2 = restore();
    if (rc || rc2 || best_k < 0) {
        b->lm_cap = cap0;
        b->passes = passes0;
        HIPCHK(hipStreamSynchronize(c->stream));
        return rc ? rc : rc2 ? rc2 : fail(CSMC_EINVAL, "no sweep form could run");
    }
    b->lm_cap = best_k < 4 ? 3 - best_k : 3;
    b->passes = best_k == 4;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (choice) *choice = best_k;
    return CSMC_OK;
}

}  // extern "C"

namespace {

// the batch-wide colour passes (csmc_sweep_passes)
// Colour passes of metropolis_constraint_adaptive! (constraint.hip): n
// sweeps of every system of the batch at the temperatures and cone widths
// already in b->T / b->sigma; accepted counts add up in cons_total.
int constraint_run(csmc_ctx *c, Batch *b, uint64_t seed, int64_t n_sweeps) {
    ConsArgs a{};
    a.P = pools(c, b);
    a.spins = b->spins.p;
    a.k0 = (uint32_t)seed;
    a.k1 = (uint32_t)(seed >> 32);
    a.ctr = b->mcount.p;
    a.T_sys = b->T.p;
    a.sigma_sys = b->sigma.p;
    a.coef = b->cons_coef.p;
    a.cnb = b->cons_nb;
    a.c0 = b->cons_c0;
    a.weight = b->cons_w;
    a.sigma0 = b->cons_sigma0;
    a.prop = b->cons_prop.p;
    a.dE = b->cons_dE.p;
    a.dC = b->cons_dC.p;
    a.u = b->cons_u.p;
    a.dec = b->cons_dec.p;
    a.lambda = b->cons_lambda.p;
    a.cval = b->cons_c.p;
    a.acc_sweep = b->cons_acc.p;
    a.acc_total = b->cons_total.p;
    int maxcol = 0, maxsites = 0;
    for (int s = 0; s < b->nsys; ++s) {
        const HostLattice &L = *c->lats[b->lat[s]];
        maxcol = std::max(maxcol, L.ncol);
        for (int col = 0; col < L.ncol; ++col)
            maxsites = std::max(maxsites, L.runs[L.colrun[col + 1] - 1][1] - L.runs[L.colrun[col]][0]);
    }
    for (int64_t sw = 0; sw < n_sweeps; ++sw) {
        HIPCHK(launch_cons_begin(b->prec, b->nsys, a, c->stream));
        for (int col = 0; col < maxcol; ++col) {
            a.colour = col;
            HIPCHK(launch_cons_pass(b->prec, b->self, b->gen, b->nsys, maxsites, a, c->stream));
        }
    }
    return CSMC_OK;
}

// the tile plans of nsw sweeps per launch of every lattice of the batch;
// false (and why) if some lattice has none
bool plan_tile_set(csmc_ctx *c, Batch *b, int nsw, size_t budget, Batch::TileSet &ts, std::string &why) {
    const size_t rec = 3 * (b->prec == 32 ? 4 : 8);
    std::vector<TileDev> td(c->lats.size(), TileDev{});
    std::vector<int32_t> tab;
    std::vector<char> done(c->lats.size(), 0);
    size_t lds = 0;
    int max_tiles = 0;
    bool flat = true;
    for (int s = 0; s < b->nsys; ++s) {
        const int id = b->lat[s];
        if (done[id]) continue;
        done[id] = 1;
        TilePlan P = plan_tiles(*c->lats[id], rec, budget, nsw);
        if (!P.ok) {
            why = "lattice " + std::to_string(id) + ": " + P.why;
            return false;
        }
        P.d.tab_off = (int32_t)tab.size();
        tab.insert(tab.end(), P.tab.begin(), P.tab.end());
        td[id] = P.d;
        lds = std::max(lds, (size_t)(P.d.nbox + 1) * rec + 4 * (size_t)P.d.p * P.d.nb);
        max_tiles = std::max(max_tiles, P.d.ntiles);
        flat = flat && P.d.L[0] == 1;
        if (c->verbose)
            fprintf(stderr,
                    "csmc: lattice %d tiled, %d sweep(s) a launch: tile %dx%dx%d cells, box %dx%dx%d x %d bases "
                    "(%.3f sites loaded, %.3f updated per site and sweep), %d tiles, %zu B LDS\n",
                    id, nsw, P.d.T[0], P.d.T[1], P.d.T[2], P.d.B[0], P.d.B[1], P.d.B[2], P.d.nb,
                    P.load_per_site, P.work_per_site, P.d.ntiles, (size_t)(P.d.nbox + 1) * rec);
    }
    if (lds > (size_t)c->max_lds) {
        why = "LDS";
        return false;
    }
    if (ts.td.alloc(td.size()) != hipSuccess || ts.tab.alloc(std::max<size_t>(tab.size(), 1)) != hipSuccess ||
        hipMemcpy(ts.td.p, td.data(), td.size() * sizeof(TileDev), hipMemcpyHostToDevice) != hipSuccess ||
        (!tab.empty() && hipMemcpy(ts.tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice) != hipSuccess)) {
        why = "device memory";
        return false;
    }
    ts.lds = lds;
    ts.max_tiles = max_tiles;
    ts.flat = flat;
    return true;
}

// the tile plans of every lattice of the batch (once per batch): one sweep
// per launch, and c->tile_sweeps per launch where every lattice allows;
// false if some lattice has no tiled form
bool tiles_ready(csmc_ctx *c, Batch *b) {
    Batch::Tiles &tl = b->tiles;
    if (tl.state != 0) return tl.state > 0;
    tl.state = -1;
    if (!plan_tile_set(c, b, 1, c->tile_lds1, tl.one, tl.why)) {
        if (c->verbose) fprintf(stderr, "csmc: batch-wide sweeps without tiles (%s)\n", tl.why.c_str());
        return false;
    }
    tl.sweeps = 1;
    std::string why;
    if (c->tile_sweeps > 1 && plan_tile_set(c, b, c->tile_sweeps, c->tile_lds, tl.many, why))
        tl.sweeps = c->tile_sweeps;
    else if (c->tile_sweeps > 1 && c->verbose)
        fprintf(stderr, "csmc: tiles of one sweep a launch (%s)\n", why.c_str());
    tl.state = 1;
    return true;
}

int sweep_passes(csmc_ctx *c, Batch *b, int32_t kind, const double *T, const double *sigma, int32_t n_sweeps,
                 uint64_t seed, double *accepted) {
    if (!valid_kind(kind)) return fail(CSMC_EINVAL, "unknown sweep kind");
    if (kind == CSMC_CONE_ADAPTIVE)
        return fail(CSMC_EINVAL, "the adaptive cone needs a host step per system (csmc_sweep_plain)");
    if (n_sweeps < 0) return fail(CSMC_EINVAL, "n_sweeps < 0");
    const bool metro = kind != CSMC_OVERRELAX && kind != CSMC_DETERMINISTIC;
    if (metro && !T) return fail(CSMC_EINVAL, "Metropolis sweeps need temperatures");
    int rc = ensure_pools(c);
    if (rc) return rc;
    std::vector<double> sg(b->nsys, 60.0);
    if (sigma) sg.assign(sigma, sigma + b->nsys);
    if (T) HIPCHK(hipMemcpyAsync(b->T.p, T, b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b->sigma.p, sg.data(), b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    // the most colours and the largest colour class of any system
    int ncol = 0, maxcol = 0;
    for (int s = 0; s < b->nsys; ++s) {
        const HostLattice &L = *c->lats[b->lat[s]];
        ncol = std::max(ncol, L.ncol);
        for (int col = 0; col < L.ncol; ++col)
            maxcol = std::max(maxcol, L.runs[L.colrun[col + 1] - 1][1] - L.runs[L.colrun[col]][0]);
    }
    PassArgs a{};
    a.P = pools(c, b);
    a.spins = b->spins.p;
    a.kind = kind;
    a.k0 = (uint32_t)seed;
    a.k1 = (uint32_t)(seed >> 32);
    a.ctr = b->mcount.p;
    a.T_sys = b->T.p;
    a.sigma_sys = b->sigma.p;
    a.acc = b->accU.p;
    a.sys0 = 0;
    HIPCHK(launch_graph_clear(b->accU.p, b->acc.p, b->nsys, c->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, c->stream));
    }
    // tiled: one launch per sweep or per tiles.sweeps sweeps, every colour
    // step of a tile from LDS
    const bool tiled = c->tile && tiles_ready(c, b);
    TileArgs ta{};
    if (tiled) {
        // the tiled sweep writes the other buffer; the pad records (never
        // written) start at zero
        if (!b->spins_alt.p || b->spins_alt.n != b->spins.n) {
            HIPCHK(b->spins_alt.alloc(b->spins.n));
            HIPCHK(hipMemsetAsync(b->spins_alt.p, 0, b->spins.n, c->stream));
        }
        ta.P = a.P;
        ta.kind = kind;
        ta.k0 = a.k0;
        ta.k1 = a.k1;
        ta.ctr = a.ctr;
        ta.T_sys = a.T_sys;
        ta.sigma_sys = a.sigma_sys;
        ta.acc = a.acc;
    }
    for (int sw = 0; sw < n_sweeps;) {
        // systems go on the grid's y dimension, at most 65535 a launch
        int done = 1;
        if (tiled) {
            const bool many = b->tiles.sweeps > 1 && sw + b->tiles.sweeps <= n_sweeps;
            const Batch::TileSet &ts = many ? b->tiles.many : b->tiles.one;
            done = many ? b->tiles.sweeps : 1;
            ta.td = ts.td.p;
            ta.tab = ts.tab.p;
            ta.flat = ts.flat ? 1 : 0;
            ta.spins_in = b->spins.p;
            ta.spins_out = b->spins_alt.p;
            for (int s0 = 0; s0 < b->nsys; s0 += 65535) {
                ta.sys0 = s0;
                HIPCHK(launch_tile(b->prec, b->self, std::min(65535, b->nsys - s0), ts.max_tiles, ts.lds, ta,
                                   c->stream));
            }
            std::swap(b->spins.p, b->spins_alt.p);  // the same size: only the roles change
        }
        for (int col = 0; !tiled && col < ncol; ++col) {  // systems with fewer colours skip the extra passes
            a.colour = col;
            for (int s0 = 0; s0 < b->nsys; s0 += 65535) {
                a.sys0 = s0;
                HIPCHK(launch_pass(b->prec, b->self, b->gen, std::min(65535, b->nsys - s0), maxcol, a, c->stream));
            }
        }
        // the Philox counter steps once per sweep (a tiled launch of several
        // sweeps draws with counter + sweep index)
        for (int k = 0; metro && k < done; ++k)
            HIPCHK(launch_graph_after_sweep(b->mcount.p, b->accU.p, b->acc.p, b->sigma.p, 0, b->nsys, false, 1.0,
                                            c->stream));
        sw += done;
    }
    if (c->timing) {
        HIPCHK(hipEventRecord(e1, c->stream));
        c->events.push_back({e0, e1});
    }
    std::vector<uint32_t> got(b->nsys);
    HIPCHK(hipMemcpyAsync(got.data(), b->accU.p, b->nsys * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (accepted)
        for (int s = 0; s < b->nsys; ++s) accepted[s] = (double)got[s];
    return CSMC_OK;
}

}  // namespace

extern "C" {

int32_t csmc_graph_sweeps(csmc_ctx *c, int32_t id, int32_t kind, const double *T, const double *sigma,
                          int32_t n_sweeps, uint64_t seed, csmc_graph **out) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!out) return fail(CSMC_EINVAL, "null graph output");
    *out = nullptr;
    if (!valid_kind(kind)) return fail(CSMC_EINVAL, "unknown sweep kind");
    if (n_sweeps < 1) return fail(CSMC_EINVAL, "n_sweeps < 1");
    const bool metro = kind != CSMC_OVERRELAX && kind != CSMC_DETERMINISTIC;
    if (metro && !T) return fail(CSMC_EINVAL, "Metropolis sweeps need temperatures");
    int rc = ensure_pools(c);
    if (rc) return rc;
    // the temperatures and cone widths the replays read
    std::vector<double> sg(b->nsys, 60.0);
    if (sigma) sg.assign(sigma, sigma + b->nsys);
    if (T) HIPCHK(hipMemcpyAsync(b->T.p, T, b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b->sigma.p, sg.data(), b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    PassArgs a{};
    a.P = pools(c, b);
    a.spins = b->spins.p;
    a.kind = kind;
    a.k0 = (uint32_t)seed;
    a.k1 = (uint32_t)(seed >> 32);
    a.ctr = b->mcount.p;  // read on the device, so every replay draws anew
    a.T_sys = b->T.p;
    a.sigma_sys = b->sigma.p;
    a.acc = b->accU.p;
    // record the same launches csmc_sweep_plain makes, plus the per-sweep
    // counter step and a clearing kernel in place of its memset
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail(CSMC_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    e = launch_graph_clear(b->accU.p, b->acc.p, b->nsys, c->stream);
    for (int s = 0; s < b->nsys && e == hipSuccess; ++s) {
        const HostLattice &L = *c->lats[b->lat[s]];
        a.sys0 = s;
        for (int sw = 0; sw < n_sweeps && e == hipSuccess; ++sw) {
            for (int col = 0; col < L.ncol && e == hipSuccess; ++col) {
                const int n_col = L.runs[L.colrun[col + 1] - 1][1] - L.runs[L.colrun[col]][0];
                a.colour = col;
                e = launch_pass(b->prec, b->self, b->gen, 1, n_col, a, c->stream);
            }
            if (metro && e == hipSuccess)
                e = launch_graph_after_sweep(b->mcount.p, b->accU.p, b->acc.p, b->sigma.p, s, 1,
                                             kind == CSMC_CONE_ADAPTIVE, (double)L.N, c->stream);
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e == hipSuccess) e = e2;
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(CSMC_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    auto *gr = new csmc_graph();
    gr->graph = g;
    gr->exec = x;
    gr->ctx = c;
    gr->batch = id;
    gr->kind = kind;
    gr->spins = b->spins.p;
    *out = gr;
    return CSMC_OK;
}

int32_t csmc_graph_launch(csmc_ctx *c, csmc_graph *g) {
    CHECK_CTX(c);
    if (!g || g->ctx != c) return fail(CSMC_EINVAL, "graph of another context");
    GET_BATCH(c, g->batch, b);
    if (b->spins.p != g->spins) {
        // tiled sweeps since the capture left the state in the batch's other
        // buffer: move it back to the one the graph was recorded with
        if (g->spins != b->spins_alt.p) return fail(CSMC_ESTATE, "the graph's spin buffer is gone");
        HIPCHK(hipMemcpyAsync(b->spins_alt.p, b->spins.p, b->spins.n, hipMemcpyDeviceToDevice, c->stream));
        std::swap(b->spins.p, b->spins_alt.p);
    }
    HIPCHK(hipGraphLaunch(g->exec, c->stream));
    return CSMC_OK;
}

int32_t csmc_graph_accepted(csmc_ctx *c, csmc_graph *g, double *accepted, double *sigma) {
    CHECK_CTX(c);
    if (!g || g->ctx != c) return fail(CSMC_EINVAL, "graph of another context");
    GET_BATCH(c, g->batch, b);
    std::vector<uint32_t> got(b->nsys);
    std::vector<double> tot(b->nsys);
    HIPCHK(hipMemcpyAsync(got.data(), b->accU.p, b->nsys * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(tot.data(), b->acc.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    if (sigma) HIPCHK(hipMemcpyAsync(sigma, b->sigma.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (accepted)
        for (int s = 0; s < b->nsys; ++s) accepted[s] = tot[s] + (double)got[s];
    return CSMC_OK;
}

int32_t csmc_graph_destroy(csmc_graph *g) {
    if (!g) return CSMC_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return CSMC_OK;
}

int32_t csmc_set_counters(csmc_ctx *c, int32_t id, const uint32_t *counters) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    HIPCHK(hipMemcpyAsync(b->mcount.p, counters, b->nsys * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_get_counters(csmc_ctx *c, int32_t id, uint32_t *counters) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    HIPCHK(hipMemcpyAsync(counters, b->mcount.p, b->nsys * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_batch_storage_mode(csmc_ctx *c, int32_t id, int32_t *mode) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!mode) return fail(CSMC_EINVAL, "null output");
    size_t bytes = 0;
    *mode = choose_lm(c, b, &bytes);
    return CSMC_OK;
}

int32_t csmc_anneal(csmc_ctx *c, int32_t id, const double *temps, int32_t n_temps,
                    int64_t t_th, int64_t or_rate, int32_t alg, double sigma0, uint64_t seed,
                    double *accepted) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (n_temps < 0 || (n_temps && !temps)) return fail(CSMC_EINVAL, "bad temperature list");
    if (alg < CSMC_METROPOLIS || alg > CSMC_CONSTRAINT_ADAPTIVE || alg == CSMC_DETERMINISTIC)
        return fail(CSMC_EINVAL, "alg must be a Metropolis kind");
    if (or_rate < 0) return fail(CSMC_EINVAL, "overrelaxation_rate < 0");
    if (alg == CSMC_CONSTRAINT_ADAPTIVE && !b->cons)
        return fail(CSMC_ESTATE, "no constraint on this batch (csmc_constraint_set)");
    if (n_temps == 0) return CSMC_OK;
    int rc = ensure_pools(c);
    if (rc) return rc;
    if (alg == CSMC_CONSTRAINT_ADAPTIVE) {
        // simulated_annealing! (monte_carlo.jl:166-187) driven from the host:
        // per temperature, sigma reset to sigma0 and t = 1 .. t_th - 1 sweep
        // steps of overrelaxation (one persistent launch per run of them) and
        // the constraint sweep every or_rate-th step
        std::vector<double> accT((size_t)b->nsys * n_temps, 0.0), got(b->nsys);
        for (int k = 0; k < n_temps; ++k) {
            std::vector<double> Tk(b->nsys, temps[k]), sg(b->nsys, sigma0);
            HIPCHK(hipMemcpyAsync(b->T.p, Tk.data(), b->nsys * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(b->sigma.p, sg.data(), b->nsys * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemsetAsync(b->cons_total.p, 0, b->nsys * 8, c->stream));
            int64_t pending_or = 0;
            auto flush_or = [&]() -> int {
                if (pending_or == 0) return CSMC_OK;
                WGArgs w = wg_base(c, b, seed);
                w.mode = WG_SWEEPS;
                w.kind = CSMC_OVERRELAX;
                w.n_sweeps = (int)pending_or;
                w.T_sys = b->T.p;
                w.acc_sys = b->acc.p;
                pending_or = 0;
                return run_wg(c, b, w);
            };
            for (int64_t t = 1; t < t_th; ++t) {
                bool metro = true;
                if (or_rate != 0) {
                    ++pending_or;
                    metro = t % or_rate == 0;
                }
                if (metro) {
                    if ((rc = flush_or())) return rc;
                    if ((rc = constraint_run(c, b, seed, 1))) return rc;
                }
            }
            if ((rc = flush_or())) return rc;
            HIPCHK(hipMemcpyAsync(got.data(), b->cons_total.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if ((rc = parts_check(b))) return rc;
            for (int s = 0; s < b->nsys; ++s) accT[(size_t)s * n_temps + k] = got[s];
        }
        if (accepted) std::copy(accT.begin(), accT.end(), accepted);
        return CSMC_OK;
    }
    DevBuf<double> dt, dacc;
    HIPCHK(dt.alloc(n_temps));
    HIPCHK(dacc.alloc((size_t)n_temps * b->nsys));
    HIPCHK(hipMemcpyAsync(dt.p, temps, n_temps * 8, hipMemcpyHostToDevice, c->stream));
    WGArgs a = wg_base(c, b, seed);
    a.mode = WG_ANNEAL;
    a.kind = alg;
    a.temps = dt.p;
    a.n_temps = n_temps;
    a.t_th = t_th;
    a.or_rate = or_rate;
    a.sigma0 = sigma0;
    a.acc_T = dacc.p;
    rc = run_wg(c, b, a);
    if (rc) return rc;
    if (accepted)
        HIPCHK(hipMemcpyAsync(accepted, dacc.p, (size_t)n_temps * b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return parts_check(b);
}

int32_t csmc_pt_init(csmc_ctx *c, int32_t id, const double *temps, const csmc_pt_params *p,
                     double sigma0, uint64_t seed) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!temps || !p) return fail(CSMC_EINVAL, "null temperatures or parameters");
    if (p->alg < CSMC_METROPOLIS || p->alg > CSMC_CONE_ADAPTIVE)
        return fail(CSMC_EINVAL, "alg must be a Metropolis kind");
    if (p->overrelaxation_rate < 0 || p->t_thermalization < 0 || p->t_measurement < 0)
        return fail(CSMC_EINVAL, "negative schedule parameter");
    if (b->nsys > 1 && p->swap_rate <= 0)
        return fail(CSMC_EINVAL, "swap_rate must be > 0 (the reference divides by it)");
    if (p->measure && p->t_measurement > 0 && p->probe_rate <= 0)
        return fail(CSMC_EINVAL, "probe_rate must be > 0 (the reference divides by it)");
    for (int s = 1; s < b->nsys; ++s)
        if (b->lat[s] != b->lat[0]) return fail(CSMC_EINVAL, "PT replicas must share one lattice");
    int rc = ensure_pools(c);
    if (rc) return rc;
    PTHost &pt = b->pt;
    if (pt.snap_pending) return fail(CSMC_ESTATE, "a save is in flight (csmc_pt_save_wait first)");
    pt.p = *p;
    pt.seed = seed;
    pt.sweep = 0;
    pt.mid = false;
    pt.corr_nk = 0;
    pt.corr_count = 0;
    pt.total = p->t_thermalization + p->t_measurement;
    pt.dosweep = p->overrelaxation_rate == 0 ? 1 : p->overrelaxation_rate;
    pt.temps.assign(temps, temps + b->nsys);
    int R = b->nsys;
    HIPCHK(pt.temps_d.alloc(R));
    HIPCHK(pt.E.alloc(R));
    HIPCHK(pt.sigma.alloc(R));
    HIPCHK(pt.acc_local.alloc(R));
    HIPCHK(pt.acc_exch.alloc(R));
    HIPCHK(pt.bins.alloc((size_t)R * 2 * BIN_DOUBLES));
    HIPCHK(pt.slot_of_sys.alloc(R));
    HIPCHK(pt.sys_of_slot.alloc(R));
    HIPCHK(pt.done.alloc(1));
    HIPCHK(pt.rec.alloc((size_t)PT_RECBUF * R * 4));
    // flags 0xFFFFFFFF: no exchange index a run reaches
    HIPCHK(hipMemsetAsync(pt.rec.p, 0xFF, (size_t)PT_RECBUF * R * 4 * 8, c->stream));
    HIPCHK(pt.err.alloc(1));
    HIPCHK(hipMemsetAsync(pt.err.p, 0, 4, c->stream));
    HIPCHK(pt.series.alloc((size_t)R * PROBE_CAP * 2));
    pt.probe_pos = 0;
    pt.chain_ok = true;
    std::vector<int32_t> ident(R);
    for (int r = 0; r < R; ++r) ident[r] = r;
    std::vector<double> sg(R, sigma0);
    HIPCHK(hipMemcpyAsync(pt.temps_d.p, temps, R * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(pt.slot_of_sys.p, ident.data(), R * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(pt.sys_of_slot.p, ident.data(), R * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(pt.sigma.p, sg.data(), R * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(pt.acc_local.p, 0, R * 8, c->stream));
    HIPCHK(hipMemsetAsync(pt.acc_exch.p, 0, R * 8, c->stream));
    HIPCHK(hipMemsetAsync(pt.bins.p, 0, (size_t)R * 2 * BIN_DOUBLES * 8, c->stream));
    HIPCHK(hipMemsetAsync(pt.done.p, 0, 4, c->stream));
    // E = total_energy(mc.lattice) (monte_carlo.jl:265)
    HIPCHK(launch_energy(b->prec, pools(c, b), b->spins.p, R, pt.E.p, nullptr, c->stream));
    rc = setup_split(c, b);
    if (rc) return rc;
    rc = setup_parts0(c, b);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    pt.init = true;
    return CSMC_OK;
}

int32_t csmc_pt_split_parts(csmc_ctx *c, int32_t id, int32_t *parts) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!b->pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    size_t bytes = 0;
    const int lm = choose_lm(c, b, &bytes);  // parts0 applies in storage mode 0 only
    if (parts) *parts = std::max(b->pt.split_G, lm == 0 ? b->pt.parts0 : 1);
    return CSMC_OK;
}

int32_t csmc_pt_advance(csmc_ctx *c, int32_t id, int64_t until, int64_t *stopped, int32_t *event) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    int rc = ensure_pools(c);
    if (rc) return rc;
    const csmc_pt_params &p = pt.p;
    const int R = b->nsys;
    if (until < 0 || until > pt.total) until = pt.total;
    auto exchange_at = [&](int64_t s) { return R > 1 && s % pt.dosweep == 0 && s % p.swap_rate == 0; };
    auto checkpoint_at = [&](int64_t s) {
        return p.checkpoint_rate > 0 && s >= p.t_thermalization && s % p.checkpoint_rate == 0;
    };
    auto report_at = [&](int64_t s1) { return p.report_interval > 0 && s1 % p.report_interval == 0; };
    // correlations are measured at the probe sweeps, after the exchange
    // (monte_carlo.jl:368-375)
    auto corr_at = [&](int64_t s) {
        return pt.corr_nk > 0 && p.measure && p.probe_rate > 0 && s >= p.t_thermalization &&
               s % p.probe_rate == 0;
    };
    // S(k) of every system, added to the sum of the temperature slot it holds
    auto corr_probe = [&]() -> int {
        HIPCHK(launch_ssf_mfma(pt.corr_plan, b->prec, b->spins.p, b->sys_d.p, pt.corr_in.perm.p,
                               pt.corr_in.tiles.p, pt.corr_in.pos.p, pt.corr_in.ks.p, b->ssf_B.p, b->ssf_P.p,
                               b->ssf_out.p, c->stream, nullptr, nullptr));
        HIPCHK(launch_ssf_accumulate(b->ssf_out.p, pt.sys_of_slot.p, R, pt.corr_nk, pt.corr_acc.p, c->stream));
        ++pt.corr_count;
        return CSMC_OK;
    };
    const int64_t max_seg = 2048;
    WGArgs a = wg_base(c, b, pt.seed);
    a.mode = WG_PT;
    a.pt.R = R;
    a.pt.alg = p.alg;
    a.pt.t_th = p.t_thermalization;
    a.pt.t_meas = p.t_measurement;
    a.pt.probe = p.probe_rate > 0 ? p.probe_rate : 1;
    a.pt.swap = p.swap_rate > 0 ? p.swap_rate : 1;
    a.pt.or_rate = p.overrelaxation_rate;
    a.pt.dosweep = pt.dosweep;
    a.pt.measure = p.measure && p.probe_rate > 0;
    a.pt.seed = pt.seed;
    a.pt.temps = pt.temps_d.p;
    a.pt.slot_of_sys = pt.slot_of_sys.p;
    a.pt.sys_of_slot = pt.sys_of_slot.p;
    a.pt.E = pt.E.p;
    a.pt.sigma = pt.sigma.p;
    a.pt.acc_local = pt.acc_local.p;
    a.pt.acc_exch = pt.acc_exch.p;
    a.pt.binners = pt.bins.p;
    a.pt.done = pt.done.p;
    a.pt.rec = pt.rec.p;
    a.pt.err = pt.err.p;
    a.spin_limit = c->spin_limit;
    a.pt.series = pt.series.p;
    a.pt.series_cap = PROBE_CAP;
    // probe sweeps of a segment (the tail of sweep se is not run when emid)
    auto probes_in = [&](int64_t sb, int64_t se, bool emid) -> int64_t {
        if (!a.pt.measure) return 0;
        const int64_t lo = std::max<int64_t>(sb, p.t_thermalization), hi = emid ? se - 1 : se;
        if (hi < lo) return 0;
        const int64_t q = a.pt.probe;
        auto fdiv = [q](int64_t x) { return x >= 0 ? x / q : -((-x + q - 1) / q); };
        return fdiv(hi) - fdiv(lo - 1);  // multiples of the probe rate in [lo, hi]
    };
    // room for n more probes in the series, folding it first if needed
    auto reserve = [&](int64_t n) -> int {
        if (pt.probe_pos + n > PROBE_CAP) {
            HIPCHK(launch_bin_fold(R, pt.series.p, PROBE_CAP, (int)pt.probe_pos, pt.bins.p, c->stream));
            pt.probe_pos = 0;
        }
        return CSMC_OK;
    };
    // Consecutive segments that end with an exchange (and no host event) and
    // are evenly spaced run as one chain: one cooperative launch, each
    // exchange decided pair by pair from published records, the spins
    // resident on chip throughout.
    const int chain_max = 1024;
    int nchain = 0;
    int64_t chain_begin = 0, chain_e0 = 0, chain_P = 0, chain_probes = 0;
    bool chain_bmid = false, chained = false;
    auto launch_one = [&](int64_t sb, int64_t se, bool bmid, bool emid, bool exch) -> int {
        const int64_t n = probes_in(sb, se, emid);
        int rc1 = reserve(n);
        if (rc1) return rc1;
        WGArgs s1 = a;
        s1.seg_begin = sb;
        s1.seg_end = se;
        s1.begin_mid = bmid;
        s1.end_mid = emid;
        s1.exchange = exch;
        s1.probe_base = pt.probe_pos;
        rc1 = run_wg(c, b, s1);
        if (rc1) return rc1;
        pt.probe_pos += n;
        return CSMC_OK;
    };
    // launch the pending chain; tail >= 0 adds the non-exchange segment from
    // the chain's last exchange to the end of sweep `tail`
    auto flush = [&](int64_t tail) -> int {
        if (nchain == 0) return CSMC_OK;
        int rc2 = CSMC_OK;
        bool done_chain = false;
        const int64_t last_e = chain_e0 + (nchain - 1) * chain_P;
        if (tail >= 0) chain_probes += probes_in(last_e, tail, false);
        if (nchain > 1 && pt.chain_ok) {
            WGArgs ca = a;
            ca.mode = WG_PT_CHAIN;
            ca.seg_begin = chain_begin;
            ca.begin_mid = chain_bmid;
            ca.chain_e0 = chain_e0;
            ca.chain_P = chain_P;
            ca.chain_tail = tail;
            ca.nseg = nchain;
            rc2 = reserve(chain_probes);
            if (rc2) return rc2;
            ca.probe_base = pt.probe_pos;
            rc2 = run_wg(c, b, ca);
            if (rc2 == CSMC_OK) {
                done_chain = chained = true;
                pt.probe_pos += chain_probes;
            } else {
                (void)hipGetLastError();  // the grid cannot be co-resident: fall back
                pt.chain_ok = false;
            }
        }
        if (!done_chain) {
            for (int g = 0; g < nchain; ++g) {
                const int64_t se = chain_e0 + g * chain_P;
                const int64_t sb = g == 0 ? chain_begin : se - chain_P;
                rc2 = launch_one(sb, se, g == 0 ? chain_bmid : true, true, true);
                if (rc2) return rc2;
            }
            if (tail >= 0) {
                rc2 = launch_one(last_e, tail, true, false, false);
                if (rc2) return rc2;
            }
        }
        nchain = 0;
        chain_probes = 0;
        return CSMC_OK;
    };
    int32_t ev = 0;
    while (true) {
        if (pt.sweep >= pt.total) {
            ev = 3;
            break;
        }
        if (pt.sweep >= until && !pt.mid) {
            ev = 0;
            break;
        }
        // plan one segment
        int64_t s = pt.sweep;
        bool begin_mid = pt.mid, end_mid = false, exch = false;
        while (true) {
            bool head = !(s == pt.sweep && begin_mid);
            if (head && (exchange_at(s) || checkpoint_at(s) || corr_at(s))) {
                end_mid = true;
                exch = exchange_at(s);
                break;
            }
            if (report_at(s + 1) || s + 1 >= pt.total || s + 1 >= until || s - pt.sweep + 1 >= max_seg) break;
            ++s;
        }
        if (c->pt_chain && R > 1 && exch && !checkpoint_at(s) && !corr_at(s)) {
            const int64_t np = probes_in(pt.sweep, s, true);
            const bool extends = nchain > 0 && chain_probes + np <= PROBE_CAP &&
                                 (nchain == 1 ? s > chain_e0 : s == chain_e0 + nchain * chain_P);
            if (nchain > 0 && !extends) {
                rc = flush(-1);
                if (rc) return rc;
            }
            if (nchain == 0) {
                chain_begin = pt.sweep;
                chain_bmid = begin_mid;
                chain_e0 = s;
            } else if (nchain == 1) {
                chain_P = s - chain_e0;
            }
            ++nchain;
            chain_probes += np;
            pt.sweep = s;
            pt.mid = true;
            if (nchain >= chain_max) {
                rc = flush(-1);
                if (rc) return rc;
            }
            continue;
        }
        if (nchain > 0 && !exch && !end_mid &&
            chain_probes + probes_in(pt.sweep, s, false) <= PROBE_CAP) {
            // the rest of the chain's last sweep (and beyond, up to a host
            // event) rides in the chain's launch
            rc = flush(s);
        } else {
            rc = flush(-1);
            if (rc) return rc;
            rc = launch_one(pt.sweep, s, begin_mid, end_mid, exch);
        }
        if (rc) return rc;
        if (end_mid) {
            pt.sweep = s;
            pt.mid = true;
            if (corr_at(s)) {
                rc = corr_probe();
                if (rc) return rc;
            }
            if (checkpoint_at(s)) {
                ev = 1;
                break;
            }
        } else {
            pt.sweep = s + 1;
            pt.mid = false;
            if (report_at(s + 1)) {
                ev = 2;
                break;
            }
        }
    }
    rc = flush(-1);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (chained || pt.parts0 > 1) {
        uint32_t err = 0;
        HIPCHK(hipMemcpy(&err, pt.err.p, 4, hipMemcpyDeviceToHost));
        if (err) return fail(CSMC_EHIP, "parallel tempering: a wait for another workgroup timed out");
    }
    if (stopped) *stopped = pt.sweep;
    if (event) *event = (pt.sweep >= pt.total && !pt.mid) ? 3 : ev;
    return CSMC_OK;
}

}  // extern "C"

namespace {

// Saved parallel-tempering state: a header, the temperatures, a table of
// section sizes and the sections, which are the device buffers that carry
// the run from one csmc_pt_advance to the next, in a fixed order.
constexpr char PT_MAGIC[8] = {'C', 'S', 'M', 'C', 'P', 'T', 'S', '1'};
constexpr int PT_NSEC = 17;

struct PtSaveHead {
    char magic[8];
    int64_t version, nsys, prec, nsec;
    int64_t sweep, mid, probe_pos, corr_nk, corr_count, split_G;
    uint64_t seed;
    csmc_pt_params params;
};

struct SecRef {
    void *p;
    size_t bytes;
};

template <class T>
SecRef sec(DevBuf<T> &d) {
    return {d.p, d.n * sizeof(T)};
}

std::vector<SecRef> pt_sections(Batch *b) {
    PTHost &pt = b->pt;
    SecRef corr = pt.corr_nk > 0 ? SecRef{pt.corr_acc.p, (size_t)b->nsys * pt.corr_nk * 9 * 8} : SecRef{nullptr, 0};
    return {sec(b->spins),     sec(b->mcount),     sec(pt.E),         sec(pt.sigma),      sec(pt.acc_local),
            sec(pt.acc_exch),  sec(pt.bins),       sec(pt.slot_of_sys), sec(pt.sys_of_slot), sec(pt.done),
            sec(pt.rec),       sec(pt.err),        sec(pt.series),    sec(pt.split_hx),   sec(pt.split_px),
            sec(pt.split_epoch), corr};
}

std::vector<unsigned char> pt_head(Batch *b) {
    PTHost &pt = b->pt;
    PtSaveHead h{};
    std::memcpy(h.magic, PT_MAGIC, 8);
    h.version = 2;  // 2: packed 12 / 24-byte spin records
    h.nsys = b->nsys;
    h.prec = b->prec;
    h.nsec = PT_NSEC;
    h.sweep = pt.sweep;
    h.mid = pt.mid;
    h.probe_pos = pt.probe_pos;
    h.corr_nk = pt.corr_nk;
    h.corr_count = pt.corr_count;
    h.split_G = pt.split_G;
    h.seed = pt.seed;
    h.params = pt.p;
    std::vector<SecRef> secs = pt_sections(b);
    std::vector<unsigned char> out(sizeof h + (size_t)b->nsys * 8 + PT_NSEC * 8);
    std::memcpy(out.data(), &h, sizeof h);
    std::memcpy(out.data() + sizeof h, pt.temps.data(), (size_t)b->nsys * 8);
    for (int k = 0; k < PT_NSEC; ++k) {
        const int64_t n = (int64_t)secs[k].bytes;
        std::memcpy(out.data() + sizeof h + (size_t)b->nsys * 8 + 8 * k, &n, 8);
    }
    return out;
}

size_t sections_bytes(const std::vector<SecRef> &secs) {
    size_t n = 0;
    for (const SecRef &r : secs) n += r.bytes;
    return n;
}

}  // namespace

extern "C" {

int32_t csmc_pt_save(csmc_ctx *c, int32_t id, void *buf, int64_t *bytes) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    if (!bytes) return fail(CSMC_EINVAL, "null size");
    std::vector<unsigned char> head = pt_head(b);
    std::vector<SecRef> secs = pt_sections(b);
    const size_t total = head.size() + sections_bytes(secs);
    if (!buf) {
        *bytes = (int64_t)total;
        return CSMC_OK;
    }
    if (*bytes < (int64_t)total) return fail(CSMC_EINVAL, "save buffer too small");
    unsigned char *o = (unsigned char *)buf;
    std::memcpy(o, head.data(), head.size());
    size_t off = head.size();
    for (const SecRef &r : secs) {
        if (r.bytes) HIPCHK(hipMemcpyAsync(o + off, r.p, r.bytes, hipMemcpyDeviceToHost, c->stream));
        off += r.bytes;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    *bytes = (int64_t)total;
    return CSMC_OK;
}

int32_t csmc_pt_save_async(csmc_ctx *c, int32_t id) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    if (pt.snap_pending) return fail(CSMC_ESTATE, "a save is in flight (csmc_pt_save_wait first)");
    if (!pt.snap_done) HIPCHK(hipEventCreateWithFlags(&pt.snap_done, hipEventDisableTiming));
    std::vector<SecRef> secs = pt_sections(b);
    const size_t total = sections_bytes(secs);
    HIPCHK(pt.snap_host.reserve(total));
    // DMA into page-locked memory, queued in order with the sweeps: the
    // launches enqueued next run after it, and the host does not wait (a
    // second stream for the copy made the following cooperative launch of
    // the chained kernel take twice as long)
    size_t off = 0;
    for (const SecRef &r : secs) {
        if (r.bytes) HIPCHK(hipMemcpyAsync(pt.snap_host.p + off, r.p, r.bytes, hipMemcpyDeviceToHost, c->stream));
        off += r.bytes;
    }
    HIPCHK(hipEventRecord(pt.snap_done, c->stream));
    pt.snap_head = pt_head(b);
    pt.snap_bytes = total;
    pt.snap_pending = true;
    return CSMC_OK;
}

int32_t csmc_pt_save_wait(csmc_ctx *c, int32_t id, void *buf, int64_t *bytes) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.snap_pending) return fail(CSMC_ESTATE, "no save in flight (csmc_pt_save_async)");
    if (!bytes) return fail(CSMC_EINVAL, "null size");
    const size_t total = pt.snap_head.size() + pt.snap_bytes;
    if (!buf) {
        *bytes = (int64_t)total;
        return CSMC_OK;
    }
    if (*bytes < (int64_t)total) return fail(CSMC_EINVAL, "save buffer too small");
    HIPCHK(hipEventSynchronize(pt.snap_done));
    std::memcpy(buf, pt.snap_head.data(), pt.snap_head.size());
    std::memcpy((unsigned char *)buf + pt.snap_head.size(), pt.snap_host.p, pt.snap_bytes);
    pt.snap_pending = false;
    *bytes = (int64_t)total;
    return CSMC_OK;
}

int32_t csmc_pt_load(csmc_ctx *c, int32_t id, const void *buf, int64_t bytes) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    if (pt.snap_pending) return fail(CSMC_ESTATE, "a save is in flight (csmc_pt_save_wait first)");
    if (!buf || bytes < (int64_t)sizeof(PtSaveHead)) return fail(CSMC_EINVAL, "saved state too short");
    const unsigned char *in = (const unsigned char *)buf;
    PtSaveHead h;
    std::memcpy(&h, in, sizeof h);
    if (std::memcmp(h.magic, PT_MAGIC, 8) != 0 || h.version != 2 || h.nsec != PT_NSEC)
        return fail(CSMC_EINVAL, "not a saved parallel-tempering state of this library version");
    // the run must be set up as the saved one was
    std::vector<unsigned char> mine = pt_head(b);
    PtSaveHead m;
    std::memcpy(&m, mine.data(), sizeof m);
    if (h.nsys != m.nsys || h.prec != m.prec)
        return fail(CSMC_EINVAL, "saved state: another number of systems or precision");
    if (h.seed != m.seed || std::memcmp(&h.params, &m.params, sizeof h.params) != 0)
        return fail(CSMC_EINVAL, "saved state: another seed or other parameters at csmc_pt_init");
    if (h.corr_nk != m.corr_nk) return fail(CSMC_EINVAL, "saved state: other correlation momenta");
    if (h.split_G != m.split_G) return fail(CSMC_EINVAL, "saved state: another split of the replicas");
    const size_t fixed = sizeof h + (size_t)h.nsys * 8 + PT_NSEC * 8;
    if ((size_t)bytes < fixed || std::memcmp(in + sizeof h, mine.data() + sizeof m, (size_t)h.nsys * 8) != 0)
        return fail(CSMC_EINVAL, "saved state: other temperatures");
    if (std::memcmp(in + sizeof h + (size_t)h.nsys * 8, mine.data() + sizeof m + (size_t)m.nsys * 8,
                    PT_NSEC * 8) != 0)
        return fail(CSMC_EINVAL, "saved state: buffers of other sizes (another lattice or batch)");
    std::vector<SecRef> secs = pt_sections(b);
    if ((size_t)bytes != fixed + sections_bytes(secs)) return fail(CSMC_EINVAL, "saved state: wrong length");
    size_t off = fixed;
    for (const SecRef &r : secs) {
        if (r.bytes) HIPCHK(hipMemcpyAsync(r.p, in + off, r.bytes, hipMemcpyHostToDevice, c->stream));
        off += r.bytes;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    pt.sweep = h.sweep;
    pt.mid = h.mid != 0;
    pt.probe_pos = h.probe_pos;
    pt.corr_count = h.corr_count;
    return CSMC_OK;
}

int32_t csmc_pt_state(csmc_ctx *c, int32_t id, int32_t *sys_of_slot, double *E, double *acc_local,
                      double *acc_exch, double *sigma) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    int R = b->nsys;
    std::vector<int32_t> sos(R);
    std::vector<double> Es(R);
    HIPCHK(hipMemcpyAsync(sos.data(), pt.sys_of_slot.p, R * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(Es.data(), pt.E.p, R * 8, hipMemcpyDeviceToHost, c->stream));
    if (acc_local) HIPCHK(hipMemcpyAsync(acc_local, pt.acc_local.p, R * 8, hipMemcpyDeviceToHost, c->stream));
    if (acc_exch) HIPCHK(hipMemcpyAsync(acc_exch, pt.acc_exch.p, R * 8, hipMemcpyDeviceToHost, c->stream));
    if (sigma) HIPCHK(hipMemcpyAsync(sigma, pt.sigma.p, R * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (sys_of_slot) std::copy(sos.begin(), sos.end(), sys_of_slot);
    if (E)
        for (int r = 0; r < R; ++r) E[r] = Es[sos[r]];
    return CSMC_OK;
}

namespace {
// log-binning readout: level l holds count, sum[2], sum2[2][2], pend[2], has
double bin_var(const double *st, int lvl, const double g[2]) {
    const double *q = st + lvl * 10;
    double n = q[0];
    if (n < 2.0) return 0.0;
    double m[2] = {q[1] / n, q[2] / n};
    double v = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double cij = (q[3 + 2 * i + j] / n - m[i] * m[j]) * n / (n - 1.0);
            v += g[i] * cij * g[j];
        }
    return v;
}
int bin_level(const double *st) {
    int l = 0;
    for (int k = 0; k < BIN_LVLS; ++k)
        if (st[k * 10] >= 32.0) l = k;
    return l;
}
double bin_err(const double *st, const double g[2]) {
    int l = bin_level(st);
    double n = st[l * 10];
    if (n <= 0) return 0.0;
    return std::sqrt(std::fabs(bin_var(st, l, g) / n));
}
}  // namespace

int32_t csmc_pt_observables(csmc_ctx *c, int32_t id, double *out) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    int R = b->nsys;
    HIPCHK(launch_bin_fold(R, pt.series.p, PROBE_CAP, (int)pt.probe_pos, pt.bins.p, c->stream));
    pt.probe_pos = 0;
    std::vector<double> bins((size_t)R * 2 * BIN_DOUBLES);
    HIPCHK(hipMemcpyAsync(bins.data(), pt.bins.p, bins.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int r = 0; r < R; ++r) {
        const double *be = &bins[(size_t)r * 2 * BIN_DOUBLES], *bm = be + BIN_DOUBLES;
        double N = (double)c->lats[b->lat[0]]->N, T = pt.temps[r];
        double ne = be[0], nm = bm[0];
        double me[2] = {ne > 0 ? be[1] / ne : 0.0, ne > 0 ? be[2] / ne : 0.0};
        double mm[2] = {nm > 0 ? bm[1] / nm : 0.0, nm > 0 ? bm[2] / nm : 0.0};
        // specific_heat / susceptibility (observables.jl:36-63)
        double C = 1.0 / (T * T) * (me[1] - me[0] * me[0]) / N;
        double gC[2] = {-2.0 * 1.0 / (T * T) * me[0] / N, 1.0 / (T * T) / N};
        double X = 1.0 / T * (mm[1] - mm[0] * mm[0]) / N;
        double gX[2] = {-2.0 * 1.0 / T * mm[0] / N, 1.0 / T / N};
        double g0[2] = {1.0, 0.0};
        double *o = out + (size_t)r * 8;
        o[0] = C;
        o[1] = bin_err(be, gC);
        o[2] = X;
        o[3] = bin_err(bm, gX);
        o[4] = me[0];
        o[5] = bin_err(be, g0);
        o[6] = mm[0];
        o[7] = bin_err(bm, g0);
    }
    return CSMC_OK;
}

}  // extern "C"

namespace {

// the positions and momenta of a correlation call on the device (3
// components each, zero-padded to the plan's sizes) and the plan; every
// system of the batch must be on one lattice.  The MFMA path visits the
// sites in the reference's order (lattice.jl:29-33, the last coordinate
// fastest), where sites 4 apart in a chunk are usually one lattice step
// apart; the rocBLAS path keeps the device order.
int ssf_prepare(csmc_ctx *c, Batch *b, const double *ks, int n_k, bool for_blas, SsfPlan *plan, SsfBufs &in) {
    if (n_k < 1 || !ks) return fail(CSMC_EINVAL, "spin correlations need momenta");
    for (int s = 1; s < b->nsys; ++s)
        if (b->lat[s] != b->lat[0])
            return fail(CSMC_EINVAL, "spin correlations need the systems of the batch to share one lattice");
    const HostLattice &L = *c->lats[b->lat[0]];
    const int N = (int)L.N, D = L.dim, R = b->nsys;
    // site positions (compute_site_positions, lattice.jl:38-51);
    // max |k| max |r| bounds every phase argument |k.r|
    std::vector<double> pref((size_t)N * D);
    export_tables(L, pref.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    double rmax = 0.0, kmax = 0.0;
    for (int i = 0; i < N; ++i) {
        double r2 = 0.0;
        for (int q = 0; q < D; ++q) r2 += pref[(size_t)i * D + q] * pref[(size_t)i * D + q];
        rmax = std::max(rmax, std::sqrt(r2));
    }
    for (int k = 0; k < n_k; ++k) {
        double k2 = 0.0;
        for (int q = 0; q < D; ++q) k2 += ks[(size_t)k * D + q] * ks[(size_t)k * D + q];
        if (!std::isfinite(k2)) return fail(CSMC_EINVAL, "momenta must be finite");
        kmax = std::max(kmax, std::sqrt(k2));
    }
    const bool fast = kmax * rmax * (1.0 + 1e-12) < SSF_FAST_ARG;
    // the step d between sites 4 apart, if every chunk of SSF_KC sites in
    // reference order has it between all its sites 4 apart (to 1e-9 of the
    // lattice's extent); then the phases rotate
    double d[3] = {0.0, 0.0, 0.0};
    bool regular = !for_blas && fast && N > 4;
    if (regular) {
        for (int q = 0; q < D; ++q) d[q] = pref[(size_t)4 * D + q] - pref[q];
        const double tol = 1e-9 * (1.0 + rmax);
        for (int i = 0; i < N && regular; ++i) {
            if (i % SSF_KC < 4) continue;  // the first 4 of a chunk come from sincos
            for (int q = 0; q < D; ++q)
                if (std::fabs(pref[(size_t)i * D + q] - pref[(size_t)(i - 4) * D + q] - d[q]) > tol) {
                    regular = false;
                    break;
                }
        }
    }
    ssf_plan(N, n_k, R, c->ncu, fast ? (regular ? 2 : 1) : 0, plan);
    for (int q = 0; q < 3; ++q) plan->d[q] = regular ? d[q] : 0.0;
    const int np = for_blas ? N : plan->npad, mp = for_blas ? n_k : plan->mpad;
    // positions in the visiting order and the momenta, zero-padded; perm[i]
    // is the device site of row i
    std::vector<double> pv((size_t)np * 3, 0.0), kv((size_t)mp * 3, 0.0);
    std::vector<int32_t> perm(N);
    for (int d0 = 0; d0 < N; ++d0) {
        const int ref = L.ref_of_dev[d0];
        const int row = for_blas ? d0 : ref;
        perm[row] = d0;
        for (int q = 0; q < D; ++q) pv[(size_t)row * 3 + q] = pref[(size_t)ref * D + q];
    }
    for (int k = 0; k < n_k; ++k)
        for (int q = 0; q < D; ++q) kv[(size_t)k * 3 + q] = ks[(size_t)k * D + q];
    std::vector<int32_t> tiles(plan->wfirst);
    tiles.insert(tiles.end(), plan->nparts.begin(), plan->nparts.end());
    HIPCHK(in.pos.reserve(pv.size()));
    HIPCHK(in.ks.reserve(kv.size()));
    HIPCHK(in.perm.reserve(perm.size()));
    HIPCHK(in.tiles.reserve(tiles.size()));
    HIPCHK(b->ssf_out.reserve((size_t)R * n_k * 9));
    HIPCHK(hipMemcpyAsync(in.pos.p, pv.data(), pv.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(in.ks.p, kv.data(), kv.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(in.perm.p, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(in.tiles.p, tiles.data(), tiles.size() * 4, hipMemcpyHostToDevice, c->stream));
    // the host vectors go out of scope: wait for the copies
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!for_blas) {
        HIPCHK(b->ssf_B.reserve((size_t)plan->npad * plan->ldb));
        HIPCHK(b->ssf_P.reserve((size_t)plan->mblocks * plan->colblocks * plan->maxparts * 2 * 64 * 16 *
                                plan->nct));
    }
    return CSMC_OK;
}

}  // namespace

extern "C" {

int32_t csmc_spin_correlations(csmc_ctx *c, int32_t id, const double *ks, int32_t n_k, double *out) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!out) return fail(CSMC_EINVAL, "spin correlations need an output");
    int rc = ensure_pools(c);
    if (rc) return rc;
    SsfPlan plan;
    rc = ssf_prepare(c, b, ks, n_k, c->ssf_blas, &plan, b->ssf_in);
    if (rc) return rc;
    const int N = plan.N, R = b->nsys;
    // with timing on, the events bracket the dominant work: the MFMA kernel,
    // or the phases and GEMMs of the rocBLAS path
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
    }
    if (c->ssf_blas) {
        if (e0) HIPCHK(hipEventRecord(e0, c->stream));
        // momenta per block: at most 2048, and the two kb x N phase matrices
        // within 2 x 512 MB
        const int kb =
            std::max(1, std::min({n_k, 2048, (int)std::max<int64_t>(64, ((int64_t)1 << 26) / N)}));
        if (!c->blas) HIPCHK(blas_create(&c->blas));
        HIPCHK(b->ssf_C.reserve((size_t)kb * N));
        HIPCHK(b->ssf_Q.reserve((size_t)kb * N));
        HIPCHK(b->ssf_B.reserve((size_t)3 * R * N));
        HIPCHK(b->ssf_Cre.reserve((size_t)3 * R * kb));
        HIPCHK(b->ssf_Cim.reserve((size_t)3 * R * kb));
        HIPCHK(launch_ssf_blas(c->blas, b->prec, b->spins.p, b->sys_d.p, R, N, b->ssf_in.pos.p, b->ssf_in.ks.p, n_k,
                               kb, b->ssf_C.p, b->ssf_Q.p, b->ssf_B.p, b->ssf_Cre.p, b->ssf_Cim.p,
                               b->ssf_out.p, c->stream));
        if (e1) HIPCHK(hipEventRecord(e1, c->stream));
    } else {
        HIPCHK(launch_ssf_mfma(plan, b->prec, b->spins.p, b->sys_d.p, b->ssf_in.perm.p, b->ssf_in.tiles.p,
                               b->ssf_in.pos.p, b->ssf_in.ks.p, b->ssf_B.p, b->ssf_P.p, b->ssf_out.p, c->stream,
                               e0, e1));
    }
    if (c->timing) c->events.push_back({e0, e1});
    HIPCHK(hipMemcpyAsync(out, b->ssf_out.p, (size_t)R * n_k * 9 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_pt_set_correlations(csmc_ctx *c, int32_t id, const double *ks, int32_t n_k) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    if (pt.sweep != 0 || pt.mid)
        return fail(CSMC_ESTATE, "correlations must be set before the first csmc_pt_advance");
    pt.corr_nk = 0;
    pt.corr_count = 0;
    if (n_k == 0) return CSMC_OK;
    int rc = ensure_pools(c);
    if (rc) return rc;
    rc = ssf_prepare(c, b, ks, n_k, false, &pt.corr_plan, pt.corr_in);
    if (rc) return rc;
    HIPCHK(pt.corr_acc.alloc((size_t)b->nsys * n_k * 9));
    HIPCHK(hipMemsetAsync(pt.corr_acc.p, 0, (size_t)b->nsys * n_k * 9 * 8, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    pt.corr_nk = n_k;
    return CSMC_OK;
}

int32_t csmc_pt_correlations(csmc_ctx *c, int32_t id, double *mean, int64_t *count) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    PTHost &pt = b->pt;
    if (!pt.init) return fail(CSMC_ESTATE, "csmc_pt_init was not called for this batch");
    if (pt.corr_nk == 0) return fail(CSMC_ESTATE, "correlations were not requested (csmc_pt_set_correlations)");
    const size_t n = (size_t)b->nsys * pt.corr_nk * 9;
    if (mean) {
        HIPCHK(hipMemcpyAsync(mean, pt.corr_acc.p, n * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        // the mean of the probes (an empty mean is NaN, as 0 / 0)
        const double inv = pt.corr_count > 0 ? 1.0 / (double)pt.corr_count : std::nan("");
        for (size_t i = 0; i < n; ++i) mean[i] = pt.corr_count > 0 ? mean[i] * inv : inv;
    }
    if (count) *count = pt.corr_count;
    return CSMC_OK;
}

int32_t csmc_constraint_set(csmc_ctx *c, int32_t id, int32_t n_basis, const double *a, const double *Q,
                            double c0, double weight, double sigma0) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (n_basis < 0 || (n_basis && (!a || !Q))) return fail(CSMC_EINVAL, "bad constraint coefficients");
    std::vector<double> coef((size_t)std::max(n_basis, 1) * 12, 0.0);
    for (int k = 0; k < n_basis; ++k) {
        for (int d = 0; d < 3; ++d) coef[k * 12 + d] = a[k * 3 + d];
        for (int d = 0; d < 9; ++d) coef[k * 12 + 3 + d] = Q[k * 9 + d];
    }
    const size_t slots = b->spins.n / b->spin_bytes();
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(b->cons_coef.alloc(coef.size()));
    HIPCHK(hipMemcpy(b->cons_coef.p, coef.data(), coef.size() * 8, hipMemcpyHostToDevice));
    if (!b->cons) {
        HIPCHK(b->cons_lambda.alloc(b->nsys));
        HIPCHK(hipMemset(b->cons_lambda.p, 0, b->nsys * 8));  // MonteCarlo: lambda = 0.0
        HIPCHK(b->cons_c.alloc(b->nsys));
        HIPCHK(b->cons_acc.alloc(b->nsys));
        HIPCHK(b->cons_total.alloc(b->nsys));
        HIPCHK(b->cons_dE.alloc(slots));
        HIPCHK(b->cons_dC.alloc(slots));
        HIPCHK(b->cons_u.alloc(slots));
        HIPCHK(b->cons_dec.alloc(slots));
        HIPCHK(b->cons_prop.alloc(slots * b->spin_bytes()));
    }
    b->cons_nb = n_basis;
    b->cons_c0 = c0;
    b->cons_w = weight;
    b->cons_sigma0 = sigma0;
    b->cons = true;
    return CSMC_OK;
}

int32_t csmc_constraint_lambda(csmc_ctx *c, int32_t id, const double *set, double *get) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!b->cons) return fail(CSMC_ESTATE, "no constraint on this batch (csmc_constraint_set)");
    if (set) HIPCHK(hipMemcpyAsync(b->cons_lambda.p, set, b->nsys * 8, hipMemcpyHostToDevice, c->stream));
    if (get) HIPCHK(hipMemcpyAsync(get, b->cons_lambda.p, b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_energy_grad_spins(csmc_ctx *c, int32_t id, double *dEdS) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!dEdS) return fail(CSMC_EINVAL, "null output");
    int rc = ensure_pools(c);
    if (rc) return rc;
    HIPCHK(launch_grad_spins(b->prec, pools(c, b), b->spins.p, b->nsys, b->maxN, b->stage.p, c->stream));
    HIPCHK(hipMemcpyAsync(dEdS, b->stage.p, (size_t)b->total * 3 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_energy_grad_couplings(csmc_ctx *c, int32_t id, double *dEdJ) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!dEdJ) return fail(CSMC_EINVAL, "null output");
    int gl = c->lats[b->lat[0]]->grad_len;
    for (int s = 1; s < b->nsys; ++s)
        if (c->lats[b->lat[s]]->grad_len != gl)
            return fail(CSMC_EINVAL, "coupling gradients need systems with one term layout");
    if (gl == 0) return CSMC_OK;
    int rc = ensure_pools(c);
    if (rc) return rc;
    DevBuf<double> d;
    HIPCHK(d.alloc((size_t)gl * b->nsys));
    HIPCHK(launch_grad_couplings(b->prec, pools(c, b), b->spins.p, b->nsys, gl, d.p, c->stream));
    HIPCHK(hipMemcpyAsync(dEdJ, d.p, (size_t)gl * b->nsys * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_energy_grad_couplings_noalloc(csmc_ctx *c, int32_t id, double *dEdJ) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    if (!dEdJ) return fail(CSMC_EINVAL, "null output");
    int gl = c->lats[b->lat[0]]->grad_len;
    for (int s = 1; s < b->nsys; ++s)
        if (c->lats[b->lat[s]]->grad_len != gl)
            return fail(CSMC_EINVAL, "coupling gradients need systems with one term layout");
    if (gl == 0) return CSMC_OK;
    const size_t bytes = (size_t)gl * b->nsys * sizeof(double);
    if (bytes > CSMC_NOALLOC_SCRATCH_BYTES)
        return fail(CSMC_EINVAL, "coupling gradients of this batch need " + std::to_string(bytes) +
                                     " bytes of scratch, more than the " +
                                     std::to_string((size_t)CSMC_NOALLOC_SCRATCH_BYTES) +
                                     " the no-allocation entry point holds; use csmc_energy_grad_couplings");
    // uploading the lattice tables would allocate: a call that uploads them
    // (csmc_set_spins, csmc_randomize_spins, any sweep) must come first
    if (c->dirty)
        return fail(CSMC_ESTATE, "lattice tables not on the device yet (csmc_set_spins or "
                                 "csmc_randomize_spins uploads them)");
    void *sp = scratch_ptr(c->device);
    if (!sp) return fail(CSMC_ESTATE, "scratch of this device not set up");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_grad_couplings(b->prec, pools(c, b), b->spins.p, b->nsys, gl, (double *)sp, c->stream));
    HIPCHK(hipMemcpyAsync(dEdJ, sp, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

int32_t csmc_gradient_descent(csmc_ctx *c, int32_t id, double eta, int32_t n_steps) {
    CHECK_CTX(c);
    GET_BATCH(c, id, b);
    int rc = ensure_pools(c);
    if (rc) return rc;
    if (n_steps < 1) return CSMC_OK;
    // the steps alternate between the spin buffer and a second one (a copy,
    // so its pad records are the same); every site is rewritten each step
    // exactly the size of the spin buffer, so the two can trade places
    if (!b->spins_alt.p || b->spins_alt.n != b->spins.n) HIPCHK(b->spins_alt.alloc(b->spins.n));
    HIPCHK(hipMemcpyAsync(b->spins_alt.p, b->spins.p, b->spins.n, hipMemcpyDeviceToDevice, c->stream));
    for (int k = 0; k < n_steps; ++k) {
        // with timing on, events bracket each step's kernel
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (c->timing) {
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            HIPCHK(hipEventRecord(e0, c->stream));
        }
        HIPCHK(launch_gd_step(b->prec, pools(c, b), b->spins.p, b->spins_alt.p, b->nsys, b->maxN, eta, c->stream));
        if (c->timing) {
            HIPCHK(hipEventRecord(e1, c->stream));
            c->events.push_back({e0, e1});
        }
        std::swap(b->spins.p, b->spins_alt.p);  // the same size: only the roles change
    }
    // the spins stay in the batch's own buffer (captured graphs hold its
    // address): after an odd number of steps, copy them back
    if (n_steps % 2) {
        std::swap(b->spins.p, b->spins_alt.p);
        HIPCHK(hipMemcpyAsync(b->spins.p, b->spins_alt.p, b->spins.n, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return CSMC_OK;
}

}  // extern "C"
