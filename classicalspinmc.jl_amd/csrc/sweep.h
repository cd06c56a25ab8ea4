// sweep.h -- the schedule of the persistent workgroup-per-system kernels,
// shared by every storage mode.
//
// wg_drive runs one launch's worth of the reference's sweep loops on the
// system the workgroup owns: n plain sweeps, the whole schedule of an
// annealing temperature list, one segment of the parallel-tempering loop, or
// a chain of segments with the pairwise exchange after each.
// The per-site work comes from a "sweeper" object with the interface
//   int  n_sites()
//   void or_sweep(), det_sweep()               one coloured sweep
//   double metro(T, kind, sigma, ctr)          one Metropolis sweep, returns
//                                              the accepted count (all threads)
//   double metro_energy(T, kind, sigma, ctr, with_m, E, m)
//                              one Metropolis sweep, then the system's FP64
//                              energy E and, with_m, summed spin m, all in
//                              one workgroup reduction; returns the accepted
//                              count (all threads)
//   void magnet(double m[3])   FP64 summed spin, the same sums as metro_energy's
//   double *red                RED_BYTES of LDS scratch
//   bool leader()              this workgroup writes the system's shared
//                              results (false for all but one part of a
//                              system shared by several workgroups)
// so each storage mode only provides its sweeps.
#pragma once
#include "device.h"
#include "kernels.h"

namespace csmc {

template <class R>
__device__ __forceinline__ const R *coef_of(const DevPools &P);
template <>
__device__ __forceinline__ const float *coef_of<float>(const DevPools &P) {
    return P.coefF;
}
template <>
__device__ __forceinline__ const double *coef_of<double>(const DevPools &P) {
    return P.coefD;
}

template <class R>
__device__ __forceinline__ spin4<double> to_d(const spin4<R> &s) {
    spin4<double> d;
    d.x = s.x;
    d.y = s.y;
    d.z = s.z;
    d.w = 0.0;
    return d;
}

__device__ __forceinline__ double combine_energy(const double p[5]) {
    // hamiltonian.jl:131
    return p[0] / 2 + p[1] / 3 + p[2] / 4 + p[3] + p[4];
}

// metropolis_adaptive! width update (metropolis.jl:129-131)
__device__ __forceinline__ double adapt_sigma(double sigma, double acc, double n) {
    double a = acc / n;
    double f = 0.5 / fmax(1.0 - a, 0.05);
    return fmin(fmax(sigma * f, 0.0), 100.0);
}

// one log-binning push (E or |M| with its square) into a slot's binner
__device__ __forceinline__ void bin_push(double *st, double v) {
    double c0 = v, c1 = v * v;
    for (int l = 0; l < BIN_LVLS; ++l) {
        double *q = st + l * 10;
        q[0] += 1.0;
        q[1] += c0;
        q[2] += c1;
        q[3] += c0 * c0;
        q[4] += c0 * c1;
        q[5] += c1 * c0;
        q[6] += c1 * c1;
        if (q[9] == 0.0) {
            q[7] = c0;
            q[8] = c1;
            q[9] = 1.0;
            return;
        }
        q[9] = 0.0;
        c0 = 0.5 * (q[7] + c0);
        c1 = 0.5 * (q[8] + c1);
    }
}

// Workgroup copy of n elements, get(i) -> put(i, v), with U independent
// loads in flight per thread (the index of each load is clamped to n - 1 so
// every load is unconditional; only the stores are guarded).
template <int U, class GET, class PUT>
__device__ __forceinline__ void wg_stage(int n, GET &&get, PUT &&put) {
    const int bd = (int)blockDim.x;
    for (int i0 = (int)threadIdx.x; i0 < n; i0 += U * bd) {
        decltype(get(0)) v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = get(min(i0 + u * bd, n - 1));
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (i0 + u * bd < n) put(i0 + u * bd, v[u]);
    }
}

// 64-bit words of (epoch << 32 | payload), each written and read as one
// relaxed device-scope atomic: a reader that sees the epoch it waits for sees
// that payload, with no fence or barrier between writer and reader.
__device__ __forceinline__ void ll_put(uint64_t *p, uint32_t v, uint32_t e) {
    __hip_atomic_store(p, ((uint64_t)e << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t ll_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the payload of word p once it carries epoch e (v: a first load of it); a
// wait that polls `limit` times, or sees another wait give up, sets *err and
// returns what it has, so every wave always reaches the end of the kernel
__device__ __forceinline__ uint32_t ll_wait(const uint64_t *p, uint64_t v, uint32_t e,
                                            uint32_t limit, uint32_t *err) {
    uint32_t n = 0;
    while ((uint32_t)(v >> 32) != e) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
        if (++n > limit) {
            __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = ll_load(p);
    }
    return (uint32_t)v;
}

__device__ __forceinline__ void ll_put_d(uint64_t *p, double v, uint32_t e) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    ll_put(p, (uint32_t)b, e);
    ll_put(p + 1, (uint32_t)(b >> 32), e);
}

__device__ __forceinline__ double ll_wait_d(const uint64_t *p, uint32_t e, uint32_t limit,
                                            uint32_t *err) {
    const uint64_t v0 = ll_load(p), v1 = ll_load(p + 1);
    const uint32_t lo = ll_wait(p, v0, e, limit, err), hi = ll_wait(p + 1, v1, e, limit, err);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The replica exchange (monte_carlo.jl:308-349) of a chained segment: the
// system in slot r publishes its energy and sigma for exchange xc = se / swap,
// waits for its partner slot's record (pairs (r, r+1) from 0 for an even xc,
// from 1 for an odd one) and decides its own pair: the same Philox draw and
// FP64 acceptance as the last-workgroup form, and both systems of a pair
// reach the same decision from the same two records.  Temperature labels
// move, spins stay.  On return slot and sigma are those of the system's new
// slot.  Only the leader workgroup of a system publishes and counts.
// scratch: 2 doubles of LDS outside the reduction area.
__device__ __forceinline__ void pair_exchange(const WGArgs &a, int64_t se, int &slot,
                                              double &sigma, double E, double *scratch,
                                              bool leader) {
    const PTDev &pt = a.pt;
    if (threadIdx.x == 0) {
        const int R = pt.R, r = slot;
        const uint32_t xc = (uint32_t)(se / pt.swap), lim = a.spin_limit >> 4;
        uint64_t *rec = pt.rec + (int64_t)(xc % PT_RECBUF) * R * 4;
        if (leader) {
            ll_put_d(rec + 4 * r, E, xc);
            ll_put_d(rec + 4 * r + 2, sigma, xc);
        }
        const int r0 = (xc % 2 == 0) ? 0 : 1;
        int lo = -1;
        if (r >= r0) lo = ((r - r0) % 2 == 0) ? r : r - 1;
        int ns = r;
        double nsig = sigma;
        if (lo >= 0 && lo + 1 < R) {
            const int hi = lo + 1, p = (r == lo) ? hi : lo;
            const double Ep = ll_wait_d(rec + 4 * p, xc, lim, pt.err);
            const double sp = ll_wait_d(rec + 4 * p + 2, xc, lim, pt.err);
            const double E_lo = (r == lo) ? E : Ep, E_hi = (r == lo) ? Ep : E;
            uint4 rr = philox4x32_10(make_uint4((uint32_t)lo, 0xFFFFFFFFu, xc, TAG_EXCH), a.k0, a.k1);
            const double u = u01d(rr.x);
            const double dbeta = 1.0 / pt.temps[hi] - 1.0 / pt.temps[lo];
            const double pacc = fmin(1.0, exp(dbeta * (E_hi - E_lo)));
            if (u < pacc) {
                ns = p;
                nsig = sp;
                if (leader) __hip_atomic_fetch_add(pt.acc_exch + r, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        scratch[0] = (double)ns;
        scratch[1] = nsig;
    }
    __syncthreads();
    slot = (int)scratch[0];
    sigma = scratch[1];
    __syncthreads();
}

template <int MODE, class SW>
__device__ __forceinline__ void wg_drive(const WGArgs &a, int sys, SW &sw) {
    uint32_t mctr = a.mcount[sys];
    auto metro_sweep = [&](double T, int kind, double &sigma) __attribute__((always_inline)) {
        double acc = sw.metro(T, kind, sigma, mctr);
        ++mctr;
        if (kind == CSMC_CONE_ADAPTIVE) sigma = adapt_sigma(sigma, acc, (double)sw.n_sites());
        return acc;
    };

    if constexpr (MODE == WG_SWEEPS) {
        const double T = a.T_sys[sys];
        double sigma = a.sigma_sys ? a.sigma_sys[sys] : 60.0;
        double acc = 0.0;
        for (int s = 0; s < a.n_sweeps; ++s) {
            if (a.kind == CSMC_OVERRELAX) sw.or_sweep();
            else if (a.kind == CSMC_DETERMINISTIC) sw.det_sweep();
            else acc += metro_sweep(T, a.kind, sigma);
        }
        if (threadIdx.x == 0 && sw.leader()) {
            if (a.acc_sys) a.acc_sys[sys] = acc;
            if (a.sigma_sys) a.sigma_sys[sys] = sigma;
        }
    } else if constexpr (MODE == WG_ANNEAL) {
        // simulated_annealing! (monte_carlo.jl:166-187), t = 1 .. t_th-1
        for (int k = 0; k < a.n_temps; ++k) {
            const double T = a.temps[k];
            double sigma = a.sigma0;
            double acc = 0.0;
            // t % or_rate kept as a running remainder: a 64-bit remainder per
            // sweep is a long instruction sequence on the GPU
            int64_t t_rem = a.or_rate != 0 ? 1 % a.or_rate : 0;
            for (int64_t t = 1; t < a.t_th; ++t) {
                if (a.or_rate != 0) {
                    sw.or_sweep();
                    if (t_rem == 0) acc += metro_sweep(T, a.kind, sigma);
                    if (++t_rem == a.or_rate) t_rem = 0;
                } else {
                    acc += metro_sweep(T, a.kind, sigma);
                }
            }
            if (threadIdx.x == 0 && sw.leader()) a.acc_T[(int64_t)sys * a.n_temps + k] = acc;
        }
    } else {
        // parallel_tempering! sweep loop (monte_carlo.jl:295-388).  State of
        // the system's current temperature slot, carried across segments.
        const PTDev &pt = a.pt;
        int slot = pt.slot_of_sys[sys];
        double T = pt.temps[slot];
        double sigma = pt.sigma[slot];
        double E = pt.E[sys];
        // probes go to the slot's series (binned later by launch_bin_fold);
        // every workgroup counts the same probe sweeps, so the index is shared
        int64_t probe_idx = a.probe_base;
        // one segment: from the post-exchange part of sweep sb (bmid) to the
        // pre-exchange part of sweep se (emid)
        // m: summed spin of the system's current state when m_now (the
        // exchange moves temperature labels, not spins, so it stays valid
        // across an exchange)
        double m[3] = {0, 0, 0};
        bool m_now = false;
        auto segment = [&](int64_t sb, int64_t se, bool bmid, bool emid) __attribute__((always_inline)) {
            double acc_local = 0.0;
            // swp % dosweep and swp % probe as running remainders (one 64-bit
            // remainder per segment instead of per sweep)
            int64_t r_do = sb % pt.dosweep, r_pr = sb % pt.probe;
            for (int64_t swp = sb; swp <= se; ++swp) {
                const bool head = !(swp == sb && bmid);
                const bool tail = !(swp == se && emid);
                const bool probe_here = pt.measure && swp >= pt.t_th && r_pr == 0;
                if (head) {
                    if (pt.or_rate != 0) {
                        sw.or_sweep();
                        m_now = false;
                    }
                    if (r_do == 0) {
                        // the Metropolis sweep and total_energy
                        // (monte_carlo.jl:303-305); the energy pass also sums
                        // the spins when this sweep has a probe
                        const bool with_m = probe_here;
                        const double acc = sw.metro_energy(T, pt.alg, sigma, mctr, with_m, E, m);
                        ++mctr;
                        if (pt.alg == CSMC_CONE_ADAPTIVE) sigma = adapt_sigma(sigma, acc, (double)sw.n_sites());
                        acc_local += acc;
                        m_now = with_m;
                    }
                }
                if (tail && probe_here) {
                    if (!m_now) {
                        sw.magnet(m);
                        m_now = true;
                    }
                    if (threadIdx.x == 0 && sw.leader()) {
                        double *q = pt.series + ((int64_t)slot * pt.series_cap + probe_idx) * 2;
                        q[0] = E;
                        q[1] = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                    }
                    ++probe_idx;
                }
                if (++r_do == pt.dosweep) r_do = 0;
                if (++r_pr == pt.probe) r_pr = 0;
            }
            // another system may hold this slot in the next segment, with no
            // ordering between the two but the exchange records: add atomically
            if (threadIdx.x == 0 && sw.leader())
                __hip_atomic_fetch_add(pt.acc_local + slot, acc_local, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        };
        if constexpr (MODE == WG_PT) {
            segment(a.seg_begin, a.seg_end, a.begin_mid, a.end_mid);
            if (threadIdx.x == 0 && sw.leader()) {
                pt.E[sys] = E;
                pt.sigma[slot] = sigma;
            }
        } else {
            // WG_PT_CHAIN: every segment ends with an exchange, decided pair
            // by pair from the partners' published records
            int64_t sb = a.seg_begin;
            bool bmid = a.begin_mid;
            for (int g = 0; g < a.nseg; ++g) {
                const int64_t se = a.chain_e0 + (int64_t)g * a.chain_P;
                segment(sb, se, bmid, true);
                pair_exchange(a, se, slot, sigma, E, sw.red + RED_XCHG, sw.leader());
                T = pt.temps[slot];
                sb = se;
                bmid = true;
            }
            if (a.chain_tail >= 0) segment(sb, a.chain_tail, bmid, false);
            if (threadIdx.x == 0 && sw.leader()) {
                pt.E[sys] = E;
                pt.sigma[slot] = sigma;
                pt.slot_of_sys[sys] = slot;
                pt.sys_of_slot[slot] = sys;
            }
        }
    }
    if (threadIdx.x == 0 && sw.leader()) a.mcount[sys] = mctr;
}

// Replica exchange (monte_carlo.jl:308-349) at the end of a PT segment, done
// by the last workgroup of the grid to finish: a completion counter, no
// waiting, so every wave reaches the end of the kernel.  Pairs (r, r+1)
// start at 0 for an even exchange index, at 1 for an odd one; the decision
// is FP64 with the Philox exchange stream; temperature labels move, spins
// stay.  Call from every thread after the workgroup's last global store;
// flag is one int of workgroup scratch in LDS.  (No static __shared__
// variable here: it would offset the dynamic LDS block and cost an address
// add on every LDS access of the sweep.)
__device__ __forceinline__ void pt_exchange_last(const WGArgs &a, int *flag) {
    // Only thread 0 fences: it wrote this system's energy and counters
    // (wg_drive), and an agent-scope fence writes back the L2, so one per
    // workgroup instead of one per wave keeps the launch tail short.  The
    // workgroup barrier then carries thread 0's acquire to the other threads.
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // release this workgroup's results
        unsigned prev = atomicAdd(a.pt.done, 1u);
        *flag = (prev == gridDim.x - 1);
        if (*flag) __threadfence();  // acquire every other workgroup's results
    }
    __syncthreads();
    if (*flag) {
        // the pairs are disjoint: one thread per pair
        const PTDev &pt = a.pt;
        const uint32_t xc = (uint32_t)(a.seg_end / pt.swap);
        const int r0 = (xc % 2 == 0) ? 0 : 1;
        for (int r = r0 + 2 * (int)threadIdx.x; r + 1 < pt.R; r += 2 * (int)blockDim.x) {
            const int p = r + 1;
            const int sr = pt.sys_of_slot[r], sp = pt.sys_of_slot[p];
            uint4 rr =
                philox4x32_10(make_uint4((uint32_t)r, 0xFFFFFFFFu, xc, TAG_EXCH), a.k0, a.k1);
            double u = u01d(rr.x);
            double dbeta = 1.0 / pt.temps[p] - 1.0 / pt.temps[r];
            double pacc = fmin(1.0, exp(dbeta * (pt.E[sp] - pt.E[sr])));
            if (u < pacc) {
                pt.sys_of_slot[r] = sp;
                pt.sys_of_slot[p] = sr;
                pt.slot_of_sys[sp] = r;
                pt.slot_of_sys[sr] = p;
                pt.acc_exch[r] += 1.0;
                pt.acc_exch[p] += 1.0;
            }
        }
        // visible to the next launch at the kernel boundary
        if (threadIdx.x == 0) *pt.done = 0u;
    }
}

}  // namespace csmc
