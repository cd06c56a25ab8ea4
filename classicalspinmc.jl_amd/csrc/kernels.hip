// kernels.hip -- gfx950 kernels of libcsmc.
//
// The sweep hot path is k_wg: one workgroup owns one system (a replica or an
// independent lattice) for a whole segment of sweeps.  The colour classes of
// the system are swept in turn with a workgroup barrier between classes, so
// no grid-wide synchronisation is ever needed.  Storage modes (LM):
//   0  spins and neighbour tables in global memory (L2 / MALL resident)
//   1  spins in LDS as three FP component arrays, neighbour tables global
//   2  spins and uint16 neighbour tables in LDS: a sweep touches no memory
//      outside the CU
// Parallel-tempering segments end with the replica exchange, done by the last
// workgroup to finish (a completion counter, no waiting).
//
// The register-table mode (neighbour offsets held in registers, spins as
// float4 in LDS) lives in kernels_reg.hip.
//
// k_pass is the plain form: one launch per colour pass, one thread per site.
#include <mutex>
#include <set>
#include <utility>

#include "sweep.h"

namespace csmc {

// interaction field split by term kind (bilinear, cubic, quartic), FP64,
// for energies and gradients
template <bool GEN, bool SP = false, class NB, class LD>
__device__ __forceinline__ void field_parts(const LatView<double> &L, const NB &nb, int b, int i,
                                            LD &&ld, double F2[3], double F3[3], double F4[3]) {
    interaction_field<double, double, false>(L, nb, b, i, ld, F2);
    F3[0] = F3[1] = F3[2] = F4[0] = F4[1] = F4[2] = 0.0;
    if (!GEN) return;
    LatView<double> Lc = L;
    Lc.K2 = Lc.K2P = 0;
    Lc.K4 = 0;
    interaction_field<double, double, true, 0, 0, SP>(Lc, nb, b, i, ld, F3);
    LatView<double> Lq = L;
    Lq.K2 = Lq.K2P = 0;
    Lq.K3 = 0;
    interaction_field<double, double, true, 0, 0, SP>(Lq, nb, b, i, ld, F4);
}

// ---------------------------------------------------------------------------
// workgroup building blocks
// ---------------------------------------------------------------------------
template <class R, class OP>
__device__ __forceinline__ void wg_colour_sweep(const LatView<R> &L, OP &&op) {
    for (int c = 0; c < L.ncol; ++c) {
        for (int r = L.colrun[c]; r < L.colrun[c + 1]; ++r) {
            int4 run = L.runs[r];
            int b = __builtin_amdgcn_readfirstlane(run.z);
            for (int i = run.x + (int)threadIdx.x; i < run.y; i += (int)blockDim.x) op(b, i);
        }
        __syncthreads();
    }
}

// Parts of one system (storage mode 0, chained PT): G workgroups share the
// system, part p taking the sites run start + (p + G k) * blockDim + thread
// of every run.  A barrier among the G workgroups ends each colour phase,
// and a sum over the system is the parts' workgroup sums added in part order,
// so every part holds the same totals.  G = 1 is the whole-system workgroup:
// a workgroup barrier, and the workgroup sum unchanged.
// Ordering among the CUs of one XCD.  The CUs of an XCD share its L2, and
// the vector L1 of a CU is write-through, so a store is visible to every CU
// of the XCD once it has completed (vmcnt) and the reader's L1 no longer
// holds the line: the code sequences of a workgroup-scope release and
// acquire when the waves of a workgroup may sit on different CUs of one XCD.
// An agent-scope fence would also write back (release) and invalidate
// (acquire) the L2 for the other XCDs.
__device__ __forceinline__ void xcd_release() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void xcd_acquire() {
    asm volatile("s_waitcnt vmcnt(0)\n\tbuffer_inv sc0" ::: "memory");
}

// the XCD (XCC) this workgroup runs on: HW_REG_XCC_ID[3:0]
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xFu;
}

struct Parts {
    int part, G, sys;
    uint32_t *bar;    // [nsys][2]: arrivals, generation
    double *psum;     // [nsys][2][G][8]
    uint32_t *err;    // set when a wait gives up
    uint32_t limit;   // polls of a wait
    int buf;          // half of psum the next sum writes
    bool xlocal;      // every part of the system runs on one XCD (found at launch)
    int gs;           // parts per system in the layout of psum and xcc (>= G)

    // every thread; the other parts' global stores before their barrier
    // call are visible after it.  With every part on one XCD the fences are
    // the XCD ones; otherwise agent-scope fences.
    __device__ void barrier() {
        if (G > 1 && xlocal) xcd_release();  // each wave: its stores are in L2
        __syncthreads();
        if (G == 1) return;
        if (threadIdx.x == 0) {
            uint32_t *cnt = bar + 2 * sys, *gen = cnt + 1;
            const uint32_t g0 = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // release this part's stores (and order the load of g0 before
            // the arrival): a release fence writes L2 back, it does not
            // invalidate it
            if (xlocal)
                xcd_release();
            else
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const uint32_t before = atomicAdd(cnt, 1u);
            if (before == (uint32_t)G - 1) {
                // the last to arrive read every arrival: acquire what the
                // others released, then open the next barrier and this one
                // (the release passes their stores on to the waiters)
                if (xlocal) {
                    xcd_acquire();
                    __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    xcd_release();  // the reset reaches L2 before the generation moves
                    __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                uint32_t n = 0;
                while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
                    if (++n > limit) {
                        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                // the other parts' stores
                if (xlocal)
                    xcd_acquire();
                else
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        __syncthreads();
    }

    // at launch, every thread of every part: publish this part's XCD, pass
    // one (agent-scope) barrier, and take the XCD fences from then on if all
    // parts of the system are on one XCD.  Every part reads the same values,
    // so all reach the same choice.
    __device__ __forceinline__ void find_xcd(uint32_t *xcc) {
        xlocal = false;
        if (G == 1) return;
        uint32_t *x = xcc + (int64_t)sys * gs;
        if (threadIdx.x == 0)
            __hip_atomic_store(x + part, xcc_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        barrier();
        bool same = true;
        const uint32_t x0 = __hip_atomic_load(x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int p = 1; p < G; ++p)
            same = same && __hip_atomic_load(x + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == x0;
        xlocal = same;
    }

    // v: this workgroup's sums (valid in every thread, as block_sum leaves
    // them); on return the system's sums, in every thread of every part
    template <int NV>
    __device__ void sum(double *v) {
        if (G == 1) return;
        uint64_t *q = reinterpret_cast<uint64_t *>(psum) + ((int64_t)(sys * 2 + buf) * gs) * 8;
        if (threadIdx.x == 0)
            for (int k = 0; k < NV; ++k)
                __hip_atomic_store(q + part * 8 + k, (uint64_t)__double_as_longlong(v[k]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        barrier();
        for (int k = 0; k < NV; ++k) {
            double t = 0.0;
            for (int p = 0; p < G; ++p)
                t += __longlong_as_double((long long)__hip_atomic_load(q + p * 8 + k, __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_AGENT));
            v[k] = t;
        }
        // a part writes this half again two sums later, after a barrier
        // that every part reaches only once it has read this one
        buf ^= 1;
    }
};

// one colour sweep of a part: op(b, i, slot) for the part's sites of each colour
// class, then the parts' barrier
template <class R, class OP>
__device__ __forceinline__ void wg_colour_sweep_p(const LatView<R> &L, Parts &pp, OP &&op) {
    const int bd = (int)blockDim.x, first = pp.part * bd + (int)threadIdx.x, step = pp.G * bd;
    // slot: the thread's visit number in the sweep x the block size + the
    // thread (the same for every sweep of a launch: see stage_part_table)
    int slot = (int)threadIdx.x;
    for (int c = 0; c < L.ncol; ++c) {
        for (int r = L.colrun[c]; r < L.colrun[c + 1]; ++r) {
            int4 run = L.runs[r];
            int b = __builtin_amdgcn_readfirstlane(run.z);
            for (int i = run.x + first; i < run.y; i += step, slot += bd) op(b, i, slot);
        }
        pp.barrier();
    }
}

// The partners of every site this thread visits in a sweep (the visit order
// of wg_colour_sweep_p) copied from the global tables into LDS as uint16
// (NbrPart); returns the table's row stride.  Every thread of the workgroup
// calls it; a barrier follows.
template <class R>
__device__ __forceinline__ int stage_part_table(const LatView<R> &L, const Parts &pp, uint16_t *tab) {
    const int bd = (int)blockDim.x, first = pp.part * bd + (int)threadIdx.x, step = pp.G * bd;
    int nvisit = 0;  // visits of the busiest thread (wave-uniform)
    for (int r = L.colrun[0]; r < L.colrun[L.ncol]; ++r) {
        const int4 run = L.runs[r];
        nvisit += (run.y - run.x + step - 1) / step;
    }
    const int stride = nvisit * bd, rows = L.K2 + 2 * L.K3 + 3 * L.K4;
    int slot = (int)threadIdx.x;
    for (int r = L.colrun[0]; r < L.colrun[L.ncol]; ++r) {
        const int4 run = L.runs[r];
        for (int i = run.x + first; i < run.y; i += step, slot += bd)
            for (int s = 0; s < rows; ++s) {
                // global row: the bilinear slots, then (after the padding
                // slots) the cubic and quartic partner rows
                const int g = s < L.K2 ? s : s + (L.K2P - L.K2);
                tab[s * stride + slot] = (uint16_t)L.nbr2[(int64_t)g * L.N + i];
            }
    }
    return stride;
}

// FP64 energy parts (bilinear, cubic, quartic, zeeman, onsite) of a system
// and its magnetisation, summed over the workgroup in one pass: out[0..4]
// the parts, out[5..7] the summed spin; valid in every thread on return
template <class R, bool GEN, bool WITH_M, class NB, class SV>
__device__ void wg_energy(const LatView<double> &Ld, const NB &nb, const SV &X, double *red,
                          double out[8]) {
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int r = Ld.colrun[0]; r < Ld.colrun[Ld.ncol]; ++r) {
        int4 run = Ld.runs[r];
        int b = __builtin_amdgcn_readfirstlane(run.z);
        const double *h = Ld.field + b * 3;
        for (int i = run.x + (int)threadIdx.x; i < run.y; i += (int)blockDim.x) {
            spin4<double> s = to_d<R>(X.ld(i));
            double F2[3], F3[3], F4[3];
            field_parts<GEN>(Ld, nb, b, i, [&](int j) { return to_d<R>(X.ld(j)); }, F2, F3, F4);
            p[0] += s.x * F2[0] + s.y * F2[1] + s.z * F2[2];
            p[1] += s.x * F3[0] + s.y * F3[1] + s.z * F3[2];
            p[2] += s.x * F4[0] + s.y * F4[1] + s.z * F4[2];
            p[3] -= s.x * h[0] + s.y * h[1] + s.z * h[2];
            p[4] += onsite_energy<double, double>(Ld, b, s.x, s.y, s.z);
            if (WITH_M) {
                p[5] += s.x;
                p[6] += s.y;
                p[7] += s.z;
            }
        }
    }
    constexpr int NV = WITH_M ? 8 : 5;
    block_sum<NV>(p, red);
    for (int k = 0; k < NV; ++k) out[k] = p[k];
}

// summed spin of a system; every thread takes the same sites, in the same
// order, as in wg_energy, so both give the same FP64 sums
template <class R, class SV>
__device__ void wg_magnet(const LatView<double> &Ld, const SV &X, double *red, double out[3],
                          int part = 0, int G = 1) {
    double m[3] = {0, 0, 0};
    const int bd = (int)blockDim.x, first = part * bd + (int)threadIdx.x, step = G * bd;
    for (int r = Ld.colrun[0]; r < Ld.colrun[Ld.ncol]; ++r) {
        const int4 run = Ld.runs[r];
        for (int i = run.x + first; i < run.y; i += step) {
            spin4<R> s = X.ld(i);
            m[0] += s.x;
            m[1] += s.y;
            m[2] += s.z;
        }
    }
    block_sum<3>(m, red);
    for (int k = 0; k < 3; ++k) out[k] = m[k];
}

// this thread's FP64 energy (the parts combined as hamiltonian.jl:131) and
// summed spin over the same sites, in the same order, as wg_energy and
// wg_magnet
// nbat(slot): the partners of the site at a visit slot (the sites and their
// order are those of wg_colour_sweep_p)
template <class R, bool GEN, bool SP, class NBAT, class SV>
__device__ __forceinline__ void thread_energy(const LatView<double> &Ld, const NBAT &nbat, const SV &X,
                                              double &e, double m[3], int part = 0, int G = 1) {
    double p[5] = {0, 0, 0, 0, 0};
    m[0] = m[1] = m[2] = 0.0;
    const int bd = (int)blockDim.x, first = part * bd + (int)threadIdx.x, step = G * bd;
    int slot = (int)threadIdx.x;
    for (int r = Ld.colrun[0]; r < Ld.colrun[Ld.ncol]; ++r) {
        int4 run = Ld.runs[r];
        int b = __builtin_amdgcn_readfirstlane(run.z);
        const double *h = Ld.field + b * 3;
        for (int i = run.x + first; i < run.y; i += step, slot += bd) {
            spin4<double> s = to_d<R>(X.ld(i));
            double F2[3], F3[3], F4[3];
            field_parts<GEN, SP>(Ld, nbat(slot), b, i, [&](int j) { return to_d<R>(X.ld(j)); }, F2, F3, F4);
            p[0] += s.x * F2[0] + s.y * F2[1] + s.z * F2[2];
            p[1] += s.x * F3[0] + s.y * F3[1] + s.z * F3[2];
            p[2] += s.x * F4[0] + s.y * F4[1] + s.z * F4[2];
            p[3] -= s.x * h[0] + s.y * h[1] + s.z * h[2];
            p[4] += onsite_energy<double, double>(Ld, b, s.x, s.y, s.z);
            m[0] += s.x;
            m[1] += s.y;
            m[2] += s.z;
        }
    }
    e = combine_energy(p);
}

// one Metropolis colour sweep; returns this thread's accepted count
template <class R, bool SELF, bool GEN, int K2C, bool SP, class NBAT, class SV>
__device__ __forceinline__ int wg_metro_local(const LatView<R> &L, const NBAT &nbat, const SV &X, R T, int kind, R sigma,
                              uint32_t k0, uint32_t k1, uint32_t stream, uint32_t ctr, Parts &pp) {
    int acc = 0;
    const R invT = R(1) / T;
    wg_colour_sweep_p<R>(L, pp, [&](int b, int i, int slot) {
        acc += metro_update<R, SELF, GEN, K2C, SP>(L, nbat(slot), X, b, i, T, invT, kind, sigma, k0, k1,
                                                   stream, ctr);
    });
    return acc;
}

template <class R, bool SELF, bool GEN, int K2C, bool SP, class NBAT, class SV>
__device__ __forceinline__ double wg_metro(const LatView<R> &L, const NBAT &nbat, const SV &X, R T, int kind, R sigma,
                           uint32_t k0, uint32_t k1, uint32_t stream, uint32_t ctr, double *red,
                           Parts &pp) {
    double v[1] = {(double)wg_metro_local<R, SELF, GEN, K2C, SP>(L, nbat, X, T, kind, sigma, k0, k1, stream,
                                                                 ctr, pp)};
    block_sum<1>(v, red);
    pp.sum<1>(v);
    return v[0];
}

// The sweeper of the global / LDS-array storage modes: sweeps walk the runs
// of each colour class with one thread per site (wg_colour_sweep_p), over
// the sites of this workgroup's part of the system.
// the LDS partner table of a workgroup's sites (NbrPart), when staged
struct PartTab {
    const uint16_t *tab;
    int stride;
};

template <class R, bool SELF, bool GEN, int K2C, bool SP, class SV, class NB, bool PTAB = false>
struct SweeperG {
    const LatView<R> &L;
    const LatView<double> &Ld;
    const SV &X;
    const NB &nb;
    double *red;
    uint32_t stream, k0, k1;
    Parts pp;
    PartTab pt;

    // the partners of the site at a visit slot: the LDS table of this
    // workgroup's sites, or the global tables
    __device__ __forceinline__ auto nbat(int slot) const {
        if constexpr (PTAB)
            return NbrPart{pt.tab, pt.stride, slot, L.K2, L.K3, L.N};
        else
            return nb;
    }
    __device__ int n_sites() const { return L.N; }
    __device__ bool leader() const { return pp.part == 0; }
    __device__ __forceinline__ void or_sweep() {
        wg_colour_sweep_p<R>(L, pp, [&](int b, int i, int slot) {
            or_update<R, GEN, K2C, SP>(L, nbat(slot), X, b, i);
        });
    }
    __device__ __forceinline__ void det_sweep() {
        wg_colour_sweep_p<R>(L, pp, [&](int b, int i, int slot) {
            det_update<R, GEN, K2C, SP>(L, nbat(slot), X, b, i);
        });
    }
    __device__ __forceinline__ double metro(double T, int kind, double sigma, uint32_t ctr) {
        auto at = [&](int slot) { return nbat(slot); };
        return wg_metro<R, SELF, GEN, K2C, SP>(L, at, X, (R)T, kind, (R)sigma, k0, k1, stream, ctr, red, pp);
    }
    __device__ __forceinline__ double metro_energy(double T, int kind, double sigma, uint32_t ctr, bool with_m,
                                   double &E, double m[3]) {
        auto at = [&](int slot) { return nbat(slot); };
        const int acc = wg_metro_local<R, SELF, GEN, K2C, SP>(L, at, X, (R)T, kind, (R)sigma, k0, k1,
                                                              stream, ctr, pp);
        double e, mm[3];
        thread_energy<R, GEN, SP>(Ld, at, X, e, mm, pp.part, pp.G);
        if (with_m) {
            double v[5] = {(double)acc, e, mm[0], mm[1], mm[2]};
            block_sum<5>(v, red);
            pp.sum<5>(v);
            E = v[1];
            m[0] = v[2];
            m[1] = v[3];
            m[2] = v[4];
            return v[0];
        }
        double v[2] = {(double)acc, e};
        block_sum<2>(v, red);
        pp.sum<2>(v);
        E = v[1];
        return v[0];
    }
    __device__ __forceinline__ void magnet(double m[3]) {
        wg_magnet<R>(Ld, X, red, m, pp.part, pp.G);
        pp.sum<3>(m);
    }
};

template <int MODE, class R, bool SELF, bool GEN, int K2C, bool SP, class SV, class NB, bool PTAB = false>
__device__ __forceinline__ void wg_run(const WGArgs &a, int sys, const LatView<R> &L,
                                       const LatView<double> &Ld, const SV &X, const NB &nb,
                                       double *red, const Parts &pp, PartTab pt = PartTab{nullptr, 0}) {
    const uint32_t stream = (uint32_t)a.P.sys[sys].stream;
    SweeperG<R, SELF, GEN, K2C, SP, SV, NB, PTAB> sw{L, Ld, X, nb, red, stream, a.k0, a.k1, pp, pt};
    wg_drive<MODE>(a, sys, sw);
}

__device__ __forceinline__ size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// ---------------------------------------------------------------------------
// k_wg: persistent workgroup-per-system sweeps.
// GEN (cubic / quartic terms) and SELF (a term names its own site) are
// launched together: the general variant is exact for every lattice, the
// fast one for bilinear-only lattices without self references.  K2C > 0
// specialises the bilinear slot count (all lattices of the batch have it).
// SP: no cubic or quartic slot of the batch is contracted in full.
// ---------------------------------------------------------------------------
// one system (or one part of it) in storage mode 0
template <class R, bool SELF, bool GEN, int MODE, int K2C, bool SP, bool PTAB = false>
__device__ __forceinline__ void wg_system0(const WGArgs &a, int sys, int part, int nparts, double *red,
                                           uint16_t *ptab = nullptr) {
    Parts pp{part, nparts, sys, a.bar0, a.psum0, a.err0, a.spin_limit, 0, false, a.parts0};
    if (nparts > 1) pp.find_xcd(a.xcc0);
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    const int4 *runs = a.P.runs + ld.run_off;
    const LatView<R> L = lat_view<R>(ld, a.P.i32, runs, coef_of<R>(a.P));
    const LatView<double> Ld = lat_view<double>(ld, a.P.i32, runs, a.P.coefD);
    if constexpr (PTAB) {
        // the partners of this workgroup's sites, staged once for the launch
        const int stride = stage_part_table<R>(L, pp, ptab);
        __syncthreads();
        wg_run<MODE, R, SELF, GEN, K2C, SP, SpinsG<R>, NbrG, true>(a, sys, L, Ld, spins_at<R>(a.spins, sd.spin_off),
                                                                  NbrG{L.nbr2, L.N}, red, pp,
                                                                  PartTab{ptab, stride});
    } else {
        wg_run<MODE, R, SELF, GEN, K2C, SP>(a, sys, L, Ld, spins_at<R>(a.spins, sd.spin_off), NbrG{L.nbr2, L.N},
                                            red, pp);
    }
}

// threads per workgroup of k_wg: 512 in FP64, 1024 in FP32, except the
// sparse general FP32 kernels (cubic / quartic lattices), 512: 256 VGPRs per
// lane instead of 128, so that their chained PT form needs no scratch
template <class R, bool SP>
constexpr int wg_threads() {
    return sizeof(R) == 8 ? 512 : (SP ? 512 : 1024);
}

int wg_threads_for(int prec, bool general, bool sparse) {
    if (prec == 64) return wg_threads<double, false>();
    return general && sparse ? wg_threads<float, true>() : wg_threads<float, false>();
}

// PTAB (storage mode 0): the partners of each workgroup's sites are staged
// in LDS for the launch (NbrPart), so a site's update gathers only spins
template <class R, int LM, bool SELF, bool GEN, int MODE, int K2C, bool SP, bool PTAB = false>
__global__ __launch_bounds__((wg_threads<R, SP>())) void k_wg(WGArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *red = reinterpret_cast<double *>(smem);  // reduction scratch

    // in storage mode 0 a system may be shared among parts0 workgroups
    const int nparts = (LM == 0 && a.parts0 > 1) ? a.parts0 : 1;
    int sys = (int)blockIdx.x / nparts, part = (int)blockIdx.x - sys * nparts;
    if (nparts > 1 && (gridDim.x / nparts) % 8 == 0) {
        // blocks go to the 8 XCDs in turn (block b to XCD b % 8): give the
        // parts of a system blocks of one XCD, so that they share its L2.
        // A bijection of blocks onto (system, part) whatever the placement.
        const int x = (int)blockIdx.x % 8, k = (int)blockIdx.x / 8;
        sys = x + 8 * (k / nparts);
        part = k % nparts;
    }

    if constexpr (LM == 0) {
        // a packed schedule (sweeps and annealing of uneven batches):
        // workgroup w runs the tasks task_off[w] .. task_off[w + 1] - 1 in
        // turn, each a system or one part of one; otherwise one task, the
        // system (part) this block stands for
        int t0 = 0, t1 = 1;
        if ((MODE == WG_SWEEPS || MODE == WG_ANNEAL) && a.tasks) {
            t0 = a.task_off[blockIdx.x];
            t1 = a.task_off[blockIdx.x + 1];
        }
        for (int t = t0; t < t1; ++t) {
            int4 tk = make_int4(sys, part, nparts, 0);
            if ((MODE == WG_SWEEPS || MODE == WG_ANNEAL) && a.tasks) tk = a.tasks[t];
            wg_system0<R, SELF, GEN, MODE, K2C, SP, PTAB>(a, tk.x, tk.y, tk.z, red,
                                                          reinterpret_cast<uint16_t *>(smem + RED_BYTES));
            __syncthreads();  // the reduction scratch is reused by the next task
        }
        if (MODE == WG_PT && a.exchange) pt_exchange_last(a, reinterpret_cast<int *>(red + RED_FLAG));
    } else {
        const Parts pp{0, 1, sys, a.bar0, a.psum0, a.err0, a.spin_limit, 0, false, 1};
        const SysDev sd = a.P.sys[sys];
        const LatDev ld = a.P.lats[sd.lat];
        const int4 *runs = a.P.runs + ld.run_off;
        const LatView<R> L = lat_view<R>(ld, a.P.i32, runs, coef_of<R>(a.P));
        const LatView<double> Ld = lat_view<double>(ld, a.P.i32, runs, a.P.coefD);
        const SpinsG<R> G = spins_at<R>(a.spins, sd.spin_off);
        const int N = L.N;
        // N + 1 entries per component: the last one is the zero pad site
        R *sx = reinterpret_cast<R *>(smem + RED_BYTES);
        SpinsSoA<R> X{sx, sx + (N + 1), sx + 2 * (N + 1)};
        wg_stage<8>(N + 1, [&](int i) { return G.ld(i); }, [&](int i, const spin4<R> &v) { X.st(i, v); });
        if constexpr (LM == 1) {
            __syncthreads();
            wg_run<MODE, R, SELF, GEN, K2C, SP>(a, sys, L, Ld, X, NbrG{L.nbr2, N}, red, pp);
        } else {
            uint16_t *idx =
                reinterpret_cast<uint16_t *>(smem + RED_BYTES + align16(3 * sizeof(R) * (N + 1)));
            const int32_t *nbr2 = L.nbr2;
            wg_stage<8>(L.K2 * N, [&](int t) { return nbr2[t]; },
                        [&](int t, int32_t v) { idx[t] = (uint16_t)v; });
            __syncthreads();
            wg_run<MODE, R, SELF, GEN, K2C, SP>(a, sys, L, Ld, X, NbrL{idx, N, L.K2}, red, pp);
        }
        // the exchange goes before the spins are stored back, so that its
        // fence writes back only the small PT state
        if (MODE == WG_PT && a.exchange) pt_exchange_last(a, reinterpret_cast<int *>(red + RED_FLAG));
        __syncthreads();
        for (int i = threadIdx.x; i < N; i += blockDim.x) G.st(i, X.ld(i));
    }
}

int max_lds_bytes() {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 65536;
    return v;
}

hipError_t allow_large_lds(const void *fn) {
    static std::mutex mu;
    static std::set<std::pair<const void *, int>> done;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    if (done.count({fn, dev})) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes());
    if (e == hipSuccess) done.insert({fn, dev});
    return e;
}

hipError_t launch_cooperative(const void *fn, int nblocks, int threads, size_t lds_bytes,
                              const WGArgs &a, hipStream_t st) {
    // the kernel's waits need every workgroup resident: refuse here, from the
    // occupancy the compiled kernel allows, before the runtime is asked
    int per_cu = 0, dev = 0, ncu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds_bytes);
    if (e != hipSuccess) return e;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if ((int64_t)per_cu * ncu < nblocks) return hipErrorCooperativeLaunchTooLarge;
    void *args[] = {const_cast<WGArgs *>(&a)};
    return hipLaunchCooperativeKernel(fn, dim3(nblocks), dim3(threads), args, (unsigned)lds_bytes, st);
}

template <class R, int LM, bool GENERAL, int MODE, int K2C, bool SP, bool PTAB = false>
static hipError_t launch_wg_t(int nsys, size_t lds_bytes, const WGArgs &a, hipStream_t st) {
    size_t bytes = RED_BYTES + ((LM || PTAB) ? lds_bytes : 0);
    const void *fn = (const void *)k_wg<R, LM, GENERAL, GENERAL, MODE, K2C, SP, PTAB>;
    if (bytes > 65536) {
        hipError_t e = allow_large_lds(fn);
        if (e != hipSuccess) return e;
    }
    const int threads = wg_threads<R, SP>();
    // the waits of a chain and the barriers of parts need every workgroup resident
    if constexpr (MODE == WG_PT_CHAIN) return launch_cooperative(fn, nsys, threads, bytes, a, st);
    if constexpr (LM == 0)
        if (a.parts0 > 1) return launch_cooperative(fn, nsys, threads, bytes, a, st);
    hipLaunchKernelGGL((k_wg<R, LM, GENERAL, GENERAL, MODE, K2C, SP, PTAB>), dim3(nsys), dim3(threads), bytes,
                       st, a);
    return hipGetLastError();
}

template <class R, int LM, bool GENERAL, int MODE>
static hipError_t launch_wg_k(int k2c, bool sparse, int nsys, size_t lds_bytes, const WGArgs &a,
                              hipStream_t st) {
    // slot-count specialisations exist for the FP32 fast path in LDS modes
    if constexpr (sizeof(R) == 4 && !GENERAL && LM > 0) {
        if (k2c == 3) return launch_wg_t<R, LM, GENERAL, MODE, 3, false>(nsys, lds_bytes, a, st);
        if (k2c == 4) return launch_wg_t<R, LM, GENERAL, MODE, 4, false>(nsys, lds_bytes, a, st);
        if (k2c == 6) return launch_wg_t<R, LM, GENERAL, MODE, 6, false>(nsys, lds_bytes, a, st);
    }
    // the sparse-pattern variant of the general FP32 kernels (1024 threads,
    // where the full contractions would spill)
    if constexpr (sizeof(R) == 4 && GENERAL) {
        // storage mode 0 with the partners of each workgroup's sites in LDS
        if constexpr (LM == 0)
            if (sparse && a.ptab) return launch_wg_t<R, LM, GENERAL, MODE, 0, true, true>(nsys, lds_bytes, a, st);
        if (sparse) return launch_wg_t<R, LM, GENERAL, MODE, 0, true>(nsys, lds_bytes, a, st);
    }
    return launch_wg_t<R, LM, GENERAL, MODE, 0, false>(nsys, lds_bytes, a, st);
}

template <class R, int LM, bool GENERAL>
static hipError_t launch_wg_g(int mode, int k2c, bool sparse, int nsys, size_t lds_bytes, const WGArgs &a,
                              hipStream_t st) {
    if (mode == WG_PT) return launch_wg_k<R, LM, GENERAL, WG_PT>(k2c, sparse, nsys, lds_bytes, a, st);
    if (mode == WG_PT_CHAIN)
        return launch_wg_k<R, LM, GENERAL, WG_PT_CHAIN>(k2c, sparse, nsys, lds_bytes, a, st);
    if (mode == WG_ANNEAL) return launch_wg_k<R, LM, GENERAL, WG_ANNEAL>(k2c, sparse, nsys, lds_bytes, a, st);
    return launch_wg_k<R, LM, GENERAL, WG_SWEEPS>(k2c, sparse, nsys, lds_bytes, a, st);
}

template <class R, int LM>
static hipError_t launch_wg_m(bool general, int k2c, bool sparse, int nsys, size_t lds_bytes,
                              const WGArgs &a, hipStream_t st) {
    return general ? launch_wg_g<R, LM, true>(a.mode, k2c, sparse, nsys, lds_bytes, a, st)
                   : launch_wg_g<R, LM, false>(a.mode, k2c, sparse, nsys, lds_bytes, a, st);
}

template <class R>
static hipError_t launch_wg_r(int lm, bool general, int k2c, bool sparse, int nsys, size_t lds_bytes,
                              const WGArgs &a, hipStream_t st) {
    if (lm == 2) return launch_wg_m<R, 2>(general, k2c, sparse, nsys, lds_bytes, a, st);
    if (lm == 1) return launch_wg_m<R, 1>(general, k2c, sparse, nsys, lds_bytes, a, st);
    return launch_wg_m<R, 0>(general, k2c, sparse, nsys, lds_bytes, a, st);
}

hipError_t launch_wg(int prec, int lm, bool self, bool gen, int k2c, bool sparse, int nsys, size_t lds_bytes,
                     const WGArgs &a, hipStream_t st) {
    const bool general = self || gen;
    if (prec == 32) return launch_wg_r<float>(lm, general, k2c, sparse, nsys, lds_bytes, a, st);
    return launch_wg_r<double>(lm, general, k2c, sparse, nsys, lds_bytes, a, st);
}

// ---------------------------------------------------------------------------
// k_pass: one colour pass, one thread per site (plain / stream form)
// ---------------------------------------------------------------------------
template <class R, bool SELF, bool GEN>
__global__ __launch_bounds__(256) void k_pass(PassArgs a) {
    const int sys = a.sys0 + blockIdx.y;
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    if (a.colour >= ld.ncol) return;
    const int4 *runs = a.P.runs + ld.run_off;
    const LatView<R> L = lat_view<R>(ld, a.P.i32, runs, coef_of<R>(a.P));
    const int start = runs[L.colrun[a.colour]].x;
    const int end = runs[L.colrun[a.colour + 1] - 1].y;
    const int i = start + blockIdx.x * blockDim.x + threadIdx.x;
    int acc = 0;
    if (i < end) {
        const SpinsG<R> X = spins_at<R>(a.spins, sd.spin_off);
        const NbrG nb{L.nbr2, L.N};
        with_wave_basis(L.sub[i], [&](int b) __attribute__((always_inline)) {
            if (a.kind == CSMC_OVERRELAX) {
                or_update<R, GEN, 0>(L, nb, X, b, i);
            } else if (a.kind == CSMC_DETERMINISTIC) {
                det_update<R, GEN, 0>(L, nb, X, b, i);
            } else {
                R sigma = a.sigma_sys ? (R)a.sigma_sys[sys] : R(60);
                R T = (R)a.T_sys[sys];
                acc = metro_update<R, SELF, GEN, 0>(L, nb, X, b, i, T, R(1) / T, a.kind, sigma, a.k0, a.k1,
                                                    (uint32_t)sd.stream, a.ctr ? a.ctr[sys] : a.ctr0);
            }
        });
    }
    if (a.kind != CSMC_OVERRELAX && a.kind != CSMC_DETERMINISTIC) {
        unsigned long long m = __ballot(acc);
        if ((threadIdx.x & 63) == 0 && m) atomicAdd(&a.acc[sys], (uint32_t)__popcll(m));
    }
}

template <class R>
static void launch_pass_r(bool self, bool gen, dim3 grid, const PassArgs &a, hipStream_t st) {
    if (gen) {
        if (self) hipLaunchKernelGGL((k_pass<R, true, true>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_pass<R, false, true>), grid, dim3(256), 0, st, a);
    } else {
        if (self) hipLaunchKernelGGL((k_pass<R, true, false>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_pass<R, false, false>), grid, dim3(256), 0, st, a);
    }
}

// after a Metropolis sweep of systems s0 .. s0 + count - 1: advance their
// Philox sweep counters, and (adaptive cone, one system) move the accepted
// count into the total and update the cone width as metropolis_adaptive!
// does (metropolis.jl:129-131), in the host path's FP64 arithmetic
__global__ void k_graph_after_sweep(uint32_t *ctr, uint32_t *acc, double *total, double *sigma, int s0,
                                    int count, int adaptive, double n_sites) {
    const int s = s0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= s0 + count) return;
    ctr[s] += 1u;
    if (adaptive) {
        const uint32_t got = acc[s];
        acc[s] = 0u;
        total[s] += (double)got;
        const double av = (double)got / n_sites;
        const double f = 0.5 / fmax(1.0 - av, 0.05);
        sigma[s] = fmin(fmax(sigma[s] * f, 0.0), 100.0);
    }
}

// clears the accepted counters and totals at the start of a replay (a fill
// kernel, not a memset node)
__global__ void k_graph_clear(uint32_t *acc, double *total, int nsys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsys) {
        acc[i] = 0u;
        total[i] = 0.0;
    }
}

hipError_t launch_graph_after_sweep(uint32_t *ctr, uint32_t *acc, double *total, double *sigma, int s0,
                                    int count, bool adaptive, double n_sites, hipStream_t st) {
    hipLaunchKernelGGL(k_graph_after_sweep, dim3((count + 255) / 256), dim3(256), 0, st, ctr, acc, total, sigma,
                       s0, count, adaptive ? 1 : 0, n_sites);
    return hipGetLastError();
}

hipError_t launch_graph_clear(uint32_t *acc, double *total, int nsys, hipStream_t st) {
    hipLaunchKernelGGL(k_graph_clear, dim3((nsys + 255) / 256), dim3(256), 0, st, acc, total, nsys);
    return hipGetLastError();
}

hipError_t launch_pass(int prec, bool self, bool gen, int nsys, int max_colour_sites,
                       const PassArgs &a, hipStream_t st) {
    dim3 grid((max_colour_sites + 255) / 256, nsys);
    if (prec == 32) launch_pass_r<float>(self, gen, grid, a, st);
    else launch_pass_r<double>(self, gen, grid, a, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// evaluation kernels (reference order outputs)
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(256) void k_site_eval(DevPools P, const void *spins, int sys,
                                                   double *H, double *e) {
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    const LatView<double> L = lat_view<double>(ld, P.i32, P.runs + ld.run_off, P.coefD);
    const SpinsG<R> S = spins_at<R>(spins, sd.spin_off);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.N) return;
    int b = L.sub[i];
    spin4<double> s = to_d<R>(S.ld(i));
    double F[3], Hl[3];
    interaction_field<double, double>(L, NbrG{L.nbr2, L.N}, b, i, [&](int j) { return to_d<R>(S.ld(j)); }, F);
    local_field<double, double>(L, b, s, F, Hl);
    int r = L.ref[i];
    if (H) {
        H[3 * (int64_t)r] = Hl[0];
        H[3 * (int64_t)r + 1] = Hl[1];
        H[3 * (int64_t)r + 2] = Hl[2];
    }
    if (e) {
        const double *h = L.field + b * 3;
        e[r] = onsite_energy<double, double>(L, b, s.x, s.y, s.z) + s.x * F[0] + s.y * F[1] +
               s.z * F[2] - (s.x * h[0] + s.y * h[1] + s.z * h[2]);
    }
}

// a few sites of one system: first store n_set spins (device indices
// set_dev, FP64 values set_val[3]), then evaluate the local field / site
// energy of n_at sites (device indices at_dev) into H[k][3] / e[k]
template <class R>
__global__ __launch_bounds__(256) void k_put_sites(DevPools P, void *spins, int sys, const int32_t *set_dev,
                                                   const double *set_val, int n_set) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_set) return;
    const SysDev sd = P.sys[sys];
    spin4<R> s;
    s.x = (R)set_val[3 * k];
    s.y = (R)set_val[3 * k + 1];
    s.z = (R)set_val[3 * k + 2];
    s.w = R(0);
    spins_at<R>(spins, sd.spin_off).st(set_dev[k], s);
}

template <class R>
__global__ __launch_bounds__(256) void k_site_eval_at(DevPools P, const void *spins, int sys, const int32_t *at_dev,
                                                      int n_at, double *H, double *e) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_at) return;
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    const LatView<double> L = lat_view<double>(ld, P.i32, P.runs + ld.run_off, P.coefD);
    const SpinsG<R> S = spins_at<R>(spins, sd.spin_off);
    const int i = at_dev[k];
    const int b = L.sub[i];
    const spin4<double> s = to_d<R>(S.ld(i));
    double F[3], Hl[3];
    interaction_field<double, double>(L, NbrG{L.nbr2, L.N}, b, i, [&](int j) { return to_d<R>(S.ld(j)); }, F);
    local_field<double, double>(L, b, s, F, Hl);
    if (H) {
        H[3 * k] = Hl[0];
        H[3 * k + 1] = Hl[1];
        H[3 * k + 2] = Hl[2];
    }
    if (e) {
        const double *h = L.field + b * 3;
        e[k] = onsite_energy<double, double>(L, b, s.x, s.y, s.z) + s.x * F[0] + s.y * F[1] + s.z * F[2] -
               (s.x * h[0] + s.y * h[1] + s.z * h[2]);
    }
}

hipError_t launch_site_eval_at(int prec, const DevPools &P, void *spins, int sys, const int32_t *set_dev,
                               const double *set_val, int n_set, const int32_t *at_dev, int n_at, double *H,
                               double *e, hipStream_t st) {
    if (n_set > 0) {
        const dim3 g((n_set + 255) / 256);
        if (prec == 32) hipLaunchKernelGGL((k_put_sites<float>), g, dim3(256), 0, st, P, spins, sys, set_dev, set_val, n_set);
        else hipLaunchKernelGGL((k_put_sites<double>), g, dim3(256), 0, st, P, spins, sys, set_dev, set_val, n_set);
    }
    if (n_at > 0) {
        const dim3 g((n_at + 255) / 256);
        if (prec == 32) hipLaunchKernelGGL((k_site_eval_at<float>), g, dim3(256), 0, st, P, spins, sys, at_dev, n_at, H, e);
        else hipLaunchKernelGGL((k_site_eval_at<double>), g, dim3(256), 0, st, P, spins, sys, at_dev, n_at, H, e);
    }
    return hipGetLastError();
}

hipError_t launch_site_eval(int prec, const DevPools &P, const void *spins, int sys, int N,
                            double *H, double *e, hipStream_t st) {
    dim3 grid((N + 255) / 256);
    if (prec == 32) hipLaunchKernelGGL((k_site_eval<float>), grid, dim3(256), 0, st, P, spins, sys, H, e);
    else hipLaunchKernelGGL((k_site_eval<double>), grid, dim3(256), 0, st, P, spins, sys, H, e);
    return hipGetLastError();
}

template <class R>
__global__ __launch_bounds__(1024) void k_energy(DevPools P, const void *spins, double *E,
                                                 double *parts) {
    __shared__ double red[BLOCK_SUM_SCRATCH];
    const int sys = blockIdx.x;
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    const LatView<double> L = lat_view<double>(ld, P.i32, P.runs + ld.run_off, P.coefD);
    const SpinsG<R> S = spins_at<R>(spins, sd.spin_off);
    double p[8];
    wg_energy<R, true, false>(L, NbrG{L.nbr2, L.N}, S, red, p);
    if (threadIdx.x == 0) {
        E[sys] = combine_energy(p);
        if (parts)
            for (int k = 0; k < 5; ++k) parts[sys * 5 + k] = p[k];
    }
}

hipError_t launch_energy(int prec, const DevPools &P, const void *spins, int nsys, double *E,
                         double *parts, hipStream_t st) {
    if (prec == 32) hipLaunchKernelGGL((k_energy<float>), dim3(nsys), dim3(1024), 0, st, P, spins, E, parts);
    else hipLaunchKernelGGL((k_energy<double>), dim3(nsys), dim3(1024), 0, st, P, spins, E, parts);
    return hipGetLastError();
}

template <class R>
__global__ __launch_bounds__(1024) void k_magnet(DevPools P, const void *spins, double *M) {
    __shared__ double red[BLOCK_SUM_SCRATCH];
    const int sys = blockIdx.x;
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    const LatView<double> L = lat_view<double>(ld, P.i32, P.runs + ld.run_off, P.coefD);
    const SpinsG<R> S = spins_at<R>(spins, sd.spin_off);
    double m[3];
    wg_magnet<R>(L, S, red, m);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) M[sys * 3 + k] = m[k];
}

hipError_t launch_magnet(int prec, const DevPools &P, const void *spins, int nsys, double *M,
                         hipStream_t st) {
    if (prec == 32) hipLaunchKernelGGL((k_magnet<float>), dim3(nsys), dim3(1024), 0, st, P, spins, M);
    else hipLaunchKernelGGL((k_magnet<double>), dim3(nsys), dim3(1024), 0, st, P, spins, M);
    return hipGetLastError();
}

// random_spin_orientation (lattice.jl:306-311) at every site; counter =
// (reference site, stream, 0, INIT)
template <class R>
__global__ __launch_bounds__(256) void k_randomize(DevPools P, void *spins, uint32_t k0,
                                                   uint32_t k1) {
    const int sys = blockIdx.y;
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ld.N) return;
    const int32_t *ref = P.i32 + ld.ref_off;
    uint4 r = philox4x32_10(make_uint4((uint32_t)ref[i], (uint32_t)sd.stream, 0u, TAG_INIT), k0, k1);
    R o[3];
    Math<R>::rand_spin((R)ld.S, Math<R>::u01(r.x), Math<R>::u01(r.y), o);
    spin4<R> s;
    s.x = o[0];
    s.y = o[1];
    s.z = o[2];
    s.w = R(0);
    spins_at<R>(spins, sd.spin_off).st(i, s);
}

hipError_t launch_randomize(int prec, const DevPools &P, void *spins, int nsys, int maxN,
                            uint32_t k0, uint32_t k1, hipStream_t st) {
    dim3 grid((maxN + 255) / 256, nsys);
    if (prec == 32) hipLaunchKernelGGL((k_randomize<float>), grid, dim3(256), 0, st, P, spins, k0, k1);
    else hipLaunchKernelGGL((k_randomize<double>), grid, dim3(256), 0, st, P, spins, k0, k1);
    return hipGetLastError();
}

template <class R>
__global__ __launch_bounds__(256) void k_scatter(DevPools P, void *spins, int sys,
                                                 const double *src) {
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ld.N) return;
    int r = P.i32[ld.ref_off + i];
    spin4<R> s;
    s.x = (R)src[3 * (int64_t)r];
    s.y = (R)src[3 * (int64_t)r + 1];
    s.z = (R)src[3 * (int64_t)r + 2];
    s.w = R(0);
    spins_at<R>(spins, sd.spin_off).st(i, s);
}

template <class R>
__global__ __launch_bounds__(256) void k_gather(DevPools P, const void *spins, int sys,
                                                double *dst) {
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ld.N) return;
    int r = P.i32[ld.ref_off + i];
    spin4<R> s = spins_at<R>(spins, sd.spin_off).ld(i);
    dst[3 * (int64_t)r] = s.x;
    dst[3 * (int64_t)r + 1] = s.y;
    dst[3 * (int64_t)r + 2] = s.z;
}

hipError_t launch_scatter_spins(int prec, const DevPools &P, void *spins, int sys, int N,
                                const double *src, hipStream_t st) {
    dim3 grid((N + 255) / 256);
    if (prec == 32) hipLaunchKernelGGL((k_scatter<float>), grid, dim3(256), 0, st, P, spins, sys, src);
    else hipLaunchKernelGGL((k_scatter<double>), grid, dim3(256), 0, st, P, spins, sys, src);
    return hipGetLastError();
}

hipError_t launch_gather_spins(int prec, const DevPools &P, const void *spins, int sys, int N,
                               double *dst, hipStream_t st) {
    dim3 grid((N + 255) / 256);
    if (prec == 32) hipLaunchKernelGGL((k_gather<float>), grid, dim3(256), 0, st, P, spins, sys, dst);
    else hipLaunchKernelGGL((k_gather<double>), grid, dim3(256), 0, st, P, spins, sys, dst);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// adjoint of total_energy (hamiltonian.jl:70-132)
// ---------------------------------------------------------------------------

// dE/ds_x by gathering over the incidence list of x: every (owner site,
// term slot, role) in which x appears contributes the derivative of that
// entry's energy, weighted 1/2, 1/3, 1/4 as in hamiltonian.jl:131; onsite
// contributes (O + O^T) s and Zeeman -h.
template <class R>
__device__ __forceinline__ void site_gradient(const DevPools &P, const LatDev &ld, const LatView<double> &L,
                                              const SpinsG<R> &S, int x, double g[3]) {
    auto sp = [&](int j) { return to_d<R>(S.ld(j)); };
    if (ld.flags & LAT_GRAD_LF) {
        // the gradient is the local field (the host checked the pairing of
        // every bond with its transpose): one gather over x's own bonds
        with_wave_basis(L.sub[x], [&](int b) __attribute__((always_inline)) {
            double F[3];
            interaction_field<double, double>(L, NbrG{L.nbr2, L.N}, b, x, sp, F);
            local_field<double, double>(L, b, sp(x), F, g);
        });
        return;
    }
    const int32_t *ioff = P.i32 + ld.inc_off;
    const Inc *inc = reinterpret_cast<const Inc *>(ioff + ld.N + 1);
    const int bx = L.sub[x];
    spin4<double> s = sp(x);
    const double *O = L.onsite + bx * 9;
    const double *h = L.field + bx * 3;
    for (int a = 0; a < 3; ++a)
        g[a] = (O[a * 3 + 0] + O[0 * 3 + a]) * s.x + (O[a * 3 + 1] + O[1 * 3 + a]) * s.y +
               (O[a * 3 + 2] + O[2 * 3 + a]) * s.z - h[a];
    for (int e = ioff[x]; e < ioff[x + 1]; ++e) {
        const Inc in = inc[e];
        const int i = in.owner, kind = in.code & 3, role = (in.code >> 2) & 3, k = in.code >> 4;
        const int b = L.sub[i];
        if (kind == 0) {
            const double *M = L.c2 + (b * L.K2P + k) * 9;
            spin4<double> si = sp(i), sj = sp(L.nbr2[(int64_t)k * L.N + i]);
            if (role == 0) {
                for (int a = 0; a < 3; ++a) g[a] += 0.5 * (M[a * 3] * sj.x + M[a * 3 + 1] * sj.y + M[a * 3 + 2] * sj.z);
            } else {
                for (int a = 0; a < 3; ++a) g[a] += 0.5 * (M[a] * si.x + M[3 + a] * si.y + M[6 + a] * si.z);
            }
        } else if (kind == 1) {
            const double *C = L.c3 + (b * L.K3 + k) * 27;
            spin4<double> v[3] = {sp(i), sp(L.nbr3[((int64_t)k * 2) * L.N + i]),
                                  sp(L.nbr3[((int64_t)k * 2 + 1) * L.N + i])};
            double w[3][3];
            for (int q = 0; q < 3; ++q) {
                w[q][0] = v[q].x;
                w[q][1] = v[q].y;
                w[q][2] = v[q].z;
            }
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int c = 0; c < 3; ++c) {
                        double cf = C[a + 3 * b2 + 9 * c] / 3.0;
                        if (role == 0) g[a] += cf * w[1][b2] * w[2][c];
                        else if (role == 1) g[b2] += cf * w[0][a] * w[2][c];
                        else g[c] += cf * w[0][a] * w[1][b2];
                    }
        } else {
            const double *Q = L.c4 + (b * L.K4 + k) * 81;
            spin4<double> v[4] = {sp(i), sp(L.nbr4[((int64_t)k * 3) * L.N + i]),
                                  sp(L.nbr4[((int64_t)k * 3 + 1) * L.N + i]),
                                  sp(L.nbr4[((int64_t)k * 3 + 2) * L.N + i])};
            double w[4][3];
            for (int q = 0; q < 4; ++q) {
                w[q][0] = v[q].x;
                w[q][1] = v[q].y;
                w[q][2] = v[q].z;
            }
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d) {
                            double cf = Q[a + 3 * b2 + 9 * c + 27 * d] / 4.0;
                            if (role == 0) g[a] += cf * w[1][b2] * w[2][c] * w[3][d];
                            else if (role == 1) g[b2] += cf * w[0][a] * w[2][c] * w[3][d];
                            else if (role == 2) g[c] += cf * w[0][a] * w[1][b2] * w[3][d];
                            else g[d] += cf * w[0][a] * w[1][b2] * w[2][c];
                        }
        }
    }
}

// Block of a one-thread-per-site kernel over a batch (a 1-D grid of bps
// blocks per system): its system and its block within the system.  Blocks go
// to the 8 XCDs in turn (block b to XCD b % 8); when the system count is a
// multiple of 8, all blocks of a system go to one XCD, so the spins its
// sites gather stay in that XCD's L2.  A bijection either way.
__device__ __forceinline__ void batch_block(int nsys, int bps, int &sys, int &blk) {
    const int b = (int)blockIdx.x;
    if (nsys % 8 == 0) {
        const int k = b / 8;
        sys = b % 8 + 8 * (k / bps);
        blk = k % bps;
    } else {
        sys = b / bps;
        blk = b % bps;
    }
}

template <class R>
__global__ __launch_bounds__(256) void k_grad_spins(DevPools P, const void *spins, double *dEdS, int nsys,
                                                    int bps) {
    int sys, blk;
    batch_block(nsys, bps, sys, blk);
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    int x = blk * blockDim.x + threadIdx.x;
    if (x >= ld.N) return;
    const LatView<double> L = lat_view<double>(ld, P.i32, P.runs + ld.run_off, P.coefD);
    const SpinsG<R> S = spins_at<R>(spins, sd.spin_off);
    double g[3];
    site_gradient<R>(P, ld, L, S, x, g);
    double *o = dEdS + 3 * (sd.site_off + L.ref[x]);
    o[0] = g[0];
    o[1] = g[1];
    o[2] = g[2];
}

hipError_t launch_grad_spins(int prec, const DevPools &P, const void *spins, int nsys, int maxN,
                             double *dEdS, hipStream_t st) {
    const int bps = (maxN + 255) / 256;
    dim3 grid((unsigned)((int64_t)bps * nsys));
    if (prec == 32) hipLaunchKernelGGL((k_grad_spins<float>), grid, dim3(256), 0, st, P, spins, dEdS, nsys, bps);
    else hipLaunchKernelGGL((k_grad_spins<double>), grid, dim3(256), 0, st, P, spins, dEdS, nsys, bps);
    return hipGetLastError();
}

// dE/dJ per unit-cell term: one workgroup per system accumulates in LDS
// with FP64 atomics.  Gradient tables at grad_off:
//   t2[nb*K2], r2[nb*K2], t3[nb*K3], r3[nb*K3], t4[nb*K4], r4[nb*K4],
//   fsrc[nb], osrc[nb], then the term offsets {field, onsite, bil, cub, qua}.
template <class R>
__global__ __launch_bounds__(1024) void k_grad_couplings(DevPools P, const void *spins,
                                                         int grad_len, double *dEdJ) {
    extern __shared__ __align__(16) unsigned char gsm[];
    double *acc = reinterpret_cast<double *>(gsm);
    const int sys = blockIdx.x;
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    const LatView<double> L = lat_view<double>(ld, P.i32, P.runs + ld.run_off, P.coefD);
    const SpinsG<R> S = spins_at<R>(spins, sd.spin_off);
    auto sp = [&](int j) { return to_d<R>(S.ld(j)); };
    for (int k = threadIdx.x; k < grad_len; k += blockDim.x) acc[k] = 0.0;
    __syncthreads();
    const int nb = L.nb;
    const int32_t *g = P.i32 + ld.grad_off;
    const int32_t *t2 = g, *r2 = t2 + nb * L.K2, *t3 = r2 + nb * L.K2, *r3 = t3 + nb * L.K3;
    const int32_t *t4 = r3 + nb * L.K3, *r4 = t4 + nb * L.K4, *fsrc = r4 + nb * L.K4;
    const int32_t *osrc = fsrc + nb, *toff = osrc + nb;
    const int off_f = toff[0], off_o = toff[1], off_2 = toff[2], off_3 = toff[3], off_4 = toff[4];
    for (int i = threadIdx.x; i < L.N; i += blockDim.x) {
        const int b = L.sub[i];
        spin4<double> s = sp(i);
        double si[3] = {s.x, s.y, s.z};
        if (fsrc[b] >= 0)
            for (int a = 0; a < 3; ++a) atomicAdd(&acc[off_f + fsrc[b] * 3 + a], -si[a]);
        if (osrc[b] >= 0)
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) atomicAdd(&acc[off_o + osrc[b] * 9 + a * 3 + c], si[a] * si[c]);
        for (int k = 0; k < L.K2; ++k) {
            int j = L.nbr2[(int64_t)k * L.N + i];
            int t = t2[b * L.K2 + k];
            if (j >= L.N || t < 0) continue;
            spin4<double> q = sp(j);
            double sj[3] = {q.x, q.y, q.z};
            bool rev = r2[b * L.K2 + k] != 0;
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) {
                    int idx = rev ? (c * 3 + a) : (a * 3 + c);
                    atomicAdd(&acc[off_2 + t * 9 + idx], 0.5 * si[a] * sj[c]);
                }
        }
        for (int k = 0; k < L.K3; ++k) {
            int j = L.nbr3[((int64_t)k * 2) * L.N + i], kk = L.nbr3[((int64_t)k * 2 + 1) * L.N + i];
            int t = t3[b * L.K3 + k];
            if (j >= L.N || t < 0) continue;
            int role = r3[b * L.K3 + k];
            spin4<double> q1 = sp(j), q2 = sp(kk);
            double w[3][3] = {{si[0], si[1], si[2]}, {q1.x, q1.y, q1.z}, {q2.x, q2.y, q2.z}};
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int c = 0; c < 3; ++c) {
                        int id[3] = {a, b2, c};
                        int tmp = id[0];
                        id[0] = id[role];
                        id[role] = tmp;
                        atomicAdd(&acc[off_3 + t * 27 + id[0] + 3 * id[1] + 9 * id[2]],
                                  w[0][a] * w[1][b2] * w[2][c] / 3.0);
                    }
        }
        for (int k = 0; k < L.K4; ++k) {
            int j = L.nbr4[((int64_t)k * 3) * L.N + i];
            int kk = L.nbr4[((int64_t)k * 3 + 1) * L.N + i];
            int l = L.nbr4[((int64_t)k * 3 + 2) * L.N + i];
            int t = t4[b * L.K4 + k];
            if (j >= L.N || t < 0) continue;
            int role = r4[b * L.K4 + k];
            spin4<double> q1 = sp(j), q2 = sp(kk), q3 = sp(l);
            double w[4][3] = {{si[0], si[1], si[2]}, {q1.x, q1.y, q1.z}, {q2.x, q2.y, q2.z},
                              {q3.x, q3.y, q3.z}};
            for (int a = 0; a < 3; ++a)
                for (int b2 = 0; b2 < 3; ++b2)
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d) {
                            int id[4] = {a, b2, c, d};
                            int tmp = id[0];
                            id[0] = id[role];
                            id[role] = tmp;
                            atomicAdd(&acc[off_4 + t * 81 + id[0] + 3 * id[1] + 9 * id[2] + 27 * id[3]],
                                      w[0][a] * w[1][b2] * w[2][c] * w[3][d] / 4.0);
                        }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < grad_len; k += blockDim.x) dEdJ[(int64_t)sys * grad_len + k] = acc[k];
}

hipError_t launch_grad_couplings(int prec, const DevPools &P, const void *spins, int nsys,
                                 int grad_len, double *dEdJ, hipStream_t st) {
    size_t bytes = (size_t)(grad_len > 0 ? grad_len : 1) * sizeof(double);
    if (bytes > 65536) {
        hipError_t e = allow_large_lds(prec == 32 ? (const void *)k_grad_couplings<float>
                                                  : (const void *)k_grad_couplings<double>);
        if (e != hipSuccess) return e;
    }
    if (prec == 32) hipLaunchKernelGGL((k_grad_couplings<float>), dim3(nsys), dim3(1024), bytes, st, P, spins, grad_len, dEdJ);
    else hipLaunchKernelGGL((k_grad_couplings<double>), dim3(nsys), dim3(1024), bytes, st, P, spins, grad_len, dEdJ);
    return hipGetLastError();
}

// Push probes [0, n) of each slot's series into its binners, in order: one
// workgroup per slot, the E binner on wave 0 and the |M| binner on wave 1,
// with the binner state staged in LDS.
__global__ __launch_bounds__(128) void k_bin_fold(const double *series, int cap, int n,
                                                   double *binners) {
    __shared__ double st[2 * BIN_DOUBLES];
    const int slot = blockIdx.x;
    double *g = binners + (int64_t)slot * 2 * BIN_DOUBLES;
    for (int k = threadIdx.x; k < 2 * BIN_DOUBLES; k += blockDim.x) st[k] = g[k];
    __syncthreads();
    if (threadIdx.x == 0 || threadIdx.x == 64) {
        const int w = threadIdx.x / 64;
        const double *q = series + (int64_t)slot * cap * 2 + w;
        for (int k = 0; k < n; ++k) bin_push(st + w * BIN_DOUBLES, q[2 * k]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * BIN_DOUBLES; k += blockDim.x) g[k] = st[k];
}

hipError_t launch_bin_fold(int R, const double *series, int cap, int n, double *binners,
                           hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_fold, dim3(R), dim3(128), 0, st, series, cap, n, binners);
    return hipGetLastError();
}

// projected gradient step on the sphere
template <class R>
__global__ __launch_bounds__(256) void k_gd(DevPools P, void *spins, const double *dEdS,
                                            double eta) {
    const int sys = blockIdx.y;
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ld.N) return;
    int r = P.i32[ld.ref_off + i];
    const SpinsG<R> S = spins_at<R>(spins, sd.spin_off);
    spin4<R> s = S.ld(i);
    const double *g = dEdS + 3 * (sd.site_off + r);
    double sx = s.x, sy = s.y, sz = s.z;
    double n = sqrt(sx * sx + sy * sy + sz * sz);
    double ux = sx / n, uy = sy / n, uz = sz / n;
    double gu = g[0] * ux + g[1] * uy + g[2] * uz;
    double px = g[0] - gu * ux, py = g[1] - gu * uy, pz = g[2] - gu * uz;
    double nx = sx - eta * px, ny = sy - eta * py, nz = sz - eta * pz;
    double nn = sqrt(nx * nx + ny * ny + nz * nz);
    s.x = (R)(nx / nn * ld.S);
    s.y = (R)(ny / nn * ld.S);
    s.z = (R)(nz / nn * ld.S);
    S.st(i, s);
}

// one projected gradient step of every site from the old spins into a
// second buffer: the gradient of site_gradient and the update of k_gd in one
// pass (the result equals k_grad_spins followed by k_gd bit for bit)
template <class R>
__global__ __launch_bounds__(256) void k_gd_step(DevPools P, const void *spins_old, void *spins_new,
                                                 double eta, int nsys, int bps) {
    int sys, blk;
    batch_block(nsys, bps, sys, blk);
    const SysDev sd = P.sys[sys];
    const LatDev ld = P.lats[sd.lat];
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= ld.N) return;
    const LatView<double> L = lat_view<double>(ld, P.i32, P.runs + ld.run_off, P.coefD);
    const SpinsG<R> S = spins_at<R>(spins_old, sd.spin_off);
    double g[3];
    site_gradient<R>(P, ld, L, S, i, g);
    spin4<R> s = S.ld(i);
    double sx = s.x, sy = s.y, sz = s.z;
    double n = sqrt(sx * sx + sy * sy + sz * sz);
    double ux = sx / n, uy = sy / n, uz = sz / n;
    double gu = g[0] * ux + g[1] * uy + g[2] * uz;
    double px = g[0] - gu * ux, py = g[1] - gu * uy, pz = g[2] - gu * uz;
    double nx = sx - eta * px, ny = sy - eta * py, nz = sz - eta * pz;
    double nn = sqrt(nx * nx + ny * ny + nz * nz);
    s.x = (R)(nx / nn * ld.S);
    s.y = (R)(ny / nn * ld.S);
    s.z = (R)(nz / nn * ld.S);
    spins_at<R>(spins_new, sd.spin_off).st(i, s);
}

hipError_t launch_gd_step(int prec, const DevPools &P, const void *spins_old, void *spins_new, int nsys, int maxN,
                          double eta, hipStream_t st) {
    const int bps = (maxN + 255) / 256;
    dim3 grid((unsigned)((int64_t)bps * nsys));
    if (prec == 32)
        hipLaunchKernelGGL((k_gd_step<float>), grid, dim3(256), 0, st, P, spins_old, spins_new, eta, nsys, bps);
    else
        hipLaunchKernelGGL((k_gd_step<double>), grid, dim3(256), 0, st, P, spins_old, spins_new, eta, nsys, bps);
    return hipGetLastError();
}

hipError_t launch_gd(int prec, const DevPools &P, void *spins, int nsys, int maxN,
                     const double *dEdS, double eta, hipStream_t st) {
    dim3 grid((maxN + 255) / 256, nsys);
    if (prec == 32) hipLaunchKernelGGL((k_gd<float>), grid, dim3(256), 0, st, P, spins, dEdS, eta);
    else hipLaunchKernelGGL((k_gd<double>), grid, dim3(256), 0, st, P, spins, dEdS, eta);
    return hipGetLastError();
}

}  // namespace csmc
