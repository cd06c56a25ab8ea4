// kernels_split.hip -- split form of the register-table chained PT kernel:
// G workgroups share one system, so a batch of R systems fills G * R compute
// units instead of R.
//
// Part g of a system holds in LDS its "local" sites: a strip of unit cells
// along the first lattice dimension (the sites it owns) and a halo around it
// deep enough for one window -- the overrelaxation sweeps and the Metropolis
// sweep between two Metropolis sweeps of the loop (monte_carlo.jl:295-305).
// Within a window the part updates, at each colour phase, a prefix of each
// colour class of its local sites: its own sites and the halo sites whose new
// values its own sites still depend on before the window ends (the host plans
// the prefixes, csrc/split.cpp).  Every update is the same arithmetic on the
// same inputs as in the whole-system kernel, the Metropolis draws are keyed by
// the reference site index, so a part's own sites evolve bit for bit as there.
//
// At the window end (after the Metropolis sweep) each part publishes the
// spins of its own sites that other parts hold (in its boundary-list order,
// as x, y and z planes, so the stores of a wave are contiguous) and its
// partial sums (accepted moves, energy parts, summed spin over its own
// sites), then reads every part's partial sums and its halo sites (the halo
// list is in the order the owners publish, so its loads are contiguous too).
// The words carry the epoch of the exchange in their upper half and are
// written and read as single 64-bit atomics, so no fence or barrier orders
// them; they are double-buffered by
// epoch parity, and since every exchange waits on every part's partial sums,
// no part can write epoch e + 2 into a buffer before all parts have read
// epoch e from it.  Each part sums the partial sums in part order, so all
// parts hold identical totals (the exchange decisions of wg_drive are then
// the same in every part; only the leader, part 0, writes shared results).
#include "reg.h"

namespace csmc {

using lds_i32 = __attribute__((address_space(3))) int32_t;
__device__ __forceinline__ int lds_int(uint32_t a) { return *(const lds_i32 *)(size_t)a; }
__device__ __forceinline__ void lds_int_st(uint32_t a, int v) { *(lds_i32 *)(size_t)a = v; }

template <int K2C, int SPT, bool DIAG>
struct SweeperS {
    const LatView<float> &L;
    const LatView<double> &Ld;
    const WGArgs &a;
    SpinsV4 X;
    double *red;
    int sys, part, ncol, nor;
    int n_loc, n_bnd, n_halo;
    const int32_t *colstart, *owncnt, *phase, *slots, *colslot, *locref;
    uint32_t base;           // LDS address of local site 0
    uint32_t bnd_l, halo_l;  // LDS addresses of the boundary list and the halo pairs
    uint32_t epoch;          // epoch of the next exchange
    int pos;                 // overrelaxation sweeps since the last window end
    uint32_t own[SPT];       // LDS address of the local site (the pad site if none)
    uint32_t nb[SPT][K2C];   // LDS address of each bilinear neighbour (pad if absent)

    struct Site {
        spin4<float> s;
        spin4<float> sj[K2C];
    };

    __device__ SweeperS(const LatView<float> &L_, const LatView<double> &Ld_, const WGArgs &a_,
                        uint32_t base_, double *red_, int sys_, int part_, const int32_t *tab,
                        const int32_t *hdr, uint32_t bnd_l_, uint32_t halo_l_)
        : L(L_), Ld(Ld_), a(a_), X(), red(red_), sys(sys_), part(part_), ncol(L_.ncol),
          nor(a_.split.nor), base(base_), bnd_l(bnd_l_), halo_l(halo_l_), pos(0) {
        n_loc = ldc(hdr + 0);
        const int nslots = ldc(hdr + 2);
        n_bnd = ldc(hdr + 3);
        n_halo = ldc(hdr + 4);
        colstart = tab + ldc(hdr + 5);
        owncnt = tab + ldc(hdr + 6);
        phase = tab + ldc(hdr + 7);
        slots = tab + ldc(hdr + 8);
        colslot = tab + ldc(hdr + 9);
        locref = tab + ldc(hdr + 11);
        const int32_t *nbl = tab + ldc(hdr + 12);
        epoch = a.split.epoch[sys];
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int qq = q < nslots ? q : nslots - 1;
            const int first = ldc(slots + 4 * qq), end = ldc(slots + 4 * qq + 1);
            const int i = first + (int)threadIdx.x;
            const bool ok = q < nslots && i < end;
            const int ic = ok ? i : first;
            own[q] = pin_address(base + 16u * (uint32_t)(ok ? i : n_loc));
#pragma unroll
            for (int k = 0; k < K2C; ++k)
                nb[q][k] = pin_address(base + 16u * (uint32_t)nbl[(int64_t)k * n_loc + ic]);
        }
    }

    __device__ int n_sites() const { return L.N; }
    __device__ bool leader() const { return part == 0; }
    __device__ int local(int q) const { return (int)((own[q] - base) >> 4); }

    __device__ __forceinline__ Site load(int q) const {
        Site v;
        v.s = X.ld(own[q]);
#pragma unroll
        for (int k = 0; k < K2C; ++k) v.sj[k] = X.ld(nb[q][k]);
        return v;
    }

    // one coloured sweep at window phase ph0 + c for colour c: the slots of
    // the colour, threads whose local site lies below the phase's prefix end
    // (the pad site of a thread without a site lies above every prefix);
    // upd(q, basis, site, couplings, owned-prefix end address)
    template <class UPD>
    __device__ __forceinline__ void colour_sweep(int ph0, UPD &&upd) {
        for (int c = 0; c < ncol; ++c) {
            const int qb = ldc(colslot + c), qe = ldc(colslot + c + 1);
            const int cb = ldc(colslot + ncol + 1 + c);
            const int cs = ldc(colstart + c);
            const uint32_t lim = base + 16u * (uint32_t)(cs + ldc(phase + ph0 + c));
            const uint32_t olim = base + 16u * (uint32_t)(cs + ldc(owncnt + c));
            BilCoef<float, K2C, DIAG> M;
            M.load(L, cb);
#pragma unroll
            for (int q = 0; q < SPT; q += 2) {
                const bool in0 = q >= qb && q < qe;
                const bool in1 = q + 1 < SPT && q + 1 >= qb && q + 1 < qe;
                if (in0 && in1) {
                    const Site v0 = load(q), v1 = load(q + 1);
                    if (own[q] < lim) upd(q, cb, v0, M, olim);
                    if (own[q + 1] < lim) upd(q + 1, cb, v1, M, olim);
                } else if (in0) {
                    if (own[q] < lim) upd(q, cb, load(q), M, olim);
                } else if (in1) {
                    if (own[q + 1] < lim) upd(q + 1, cb, load(q + 1), M, olim);
                }
            }
            __syncthreads();
        }
    }

    __device__ void or_sweep() {
        colour_sweep(pos * ncol, [&](int q, int b, const Site &v, const BilCoef<float, K2C, DIAG> &M,
                                     uint32_t) {
            spin4<float> n;
            float F[3], H[3];
            M.field(v.sj, F);
            local_field<float, float>(L, b, v.s, F, H);
            const bool zero = H[0] == 0.0f && H[1] == 0.0f && H[2] == 0.0f;
            const float hh = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
            const float proj = Math<float>::div(2.0f * (v.s.x * H[0] + v.s.y * H[1] + v.s.z * H[2]), hh);
            n.x = zero ? v.s.x : -v.s.x + proj * H[0];
            n.y = zero ? v.s.y : -v.s.y + proj * H[1];
            n.z = zero ? v.s.z : -v.s.z + proj * H[2];
            n.w = 0.0f;
            X.st(own[q], n);
        });
        // a window has nor overrelaxation sweeps; a launch that starts inside
        // one counts from 0, which only updates larger prefixes than needed
        if (pos < nor - 1) ++pos;
    }

    __device__ void det_sweep() {}

    __device__ double metro(double, int, double, uint32_t) { return 0.0; }  // PT only

    // The Metropolis sweep that ends the window, then the window exchange,
    // which also sums the accepted count, energy and spin of the system.
    __device__ double metro_energy(double T, int kind, double sigma, uint32_t ctr, bool with_m,
                                   double &E, double m[3]) {
        const float Tf = (float)T, invT = 1.0f / Tf, sg = (float)sigma;
        int acc = 0;
        const uint32_t stream = (uint32_t)a.P.sys[sys].stream;
        colour_sweep(nor * ncol, [&](int q, int b, const Site &v, const BilCoef<float, K2C, DIAG> &M,
                                     uint32_t olim) {
            const uint32_t ref = (uint32_t)locref[local(q)];
            uint4 r = philox4x32_10(make_uint4(ref, stream, ctr, TAG_METRO), a.k0, a.k1);
            spin4<float> n = metro_propose<float>(L, v.s, r, kind, sg);
            float F[3];
            M.field(v.sj, F);
            int ac = metro_accept<float>(metro_dE<float>(L, b, v.s, n, F), r, Tf, invT);
            metro_commit<float>(L, X, own[q], ac, v.s, n);
            if (own[q] < olim) acc += ac;
        });
        pos = 0;
        double p[SPLIT_NV];
        own_sums(p, true);
        p[0] = (double)acc;
        publish_own(p);
        exchange(true);
        E = red[RED_TOT + 1];
        if (with_m)
            for (int k = 0; k < 3; ++k) m[k] = red[RED_TOT + 2 + k];
        return red[RED_TOT];
    }

    __device__ void magnet(double out[3]) {
        double p[SPLIT_NV];
        own_sums(p, false);
        publish_own(p);
        exchange(false);
        for (int k = 0; k < 3; ++k) out[k] = red[RED_TOT + 2 + k];
    }

    // per-thread sums over this part's own sites: p[1] the FP64 energy
    // (bilinear / 2 + Zeeman + onsite) when with_e, p[2..4] summed spin;
    // p[0] = 0
    __device__ void own_sums(double p[SPLIT_NV], bool with_e) {
        for (int k = 0; k < SPLIT_NV; ++k) p[k] = 0.0;
        double bil = 0.0, zee = 0.0, ons = 0.0;
        for (int c = 0; c < ncol; ++c) {
            const int qb = ldc(colslot + c), qe = ldc(colslot + c + 1);
            const int b = ldc(colslot + ncol + 1 + c);
            const uint32_t olim = base + 16u * (uint32_t)(ldc(colstart + c) + ldc(owncnt + c));
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                if (q >= qb && q < qe && own[q] < olim) {
                    if (with_e) {
                        const Site v = load(q);
                        const double sx = v.s.x, sy = v.s.y, sz = v.s.z;
                        BilCoef<double, K2C, DIAG> M;
                        M.load(Ld, b);
                        double F[3];
                        M.field(v.sj, F);
                        const double *h = Ld.field + b * 3;
                        bil += sx * F[0] + sy * F[1] + sz * F[2];
                        zee -= sx * ldc(h) + sy * ldc(h + 1) + sz * ldc(h + 2);
                        ons += onsite_energy<double, double>(Ld, b, sx, sy, sz);
                        p[2] += sx;
                        p[3] += sy;
                        p[4] += sz;
                    } else {
                        const spin4<float> s = X.ld(own[q]);
                        p[2] += s.x;
                        p[3] += s.y;
                        p[4] += s.z;
                    }
                }
            }
        }
        if (with_e) p[1] = bil / 2 + zee + ons;
    }

    // workgroup sums of p into RED_OWN
    __device__ void publish_own(double p[SPLIT_NV]) {
        block_sum<SPLIT_NV>(p, red);
        if (threadIdx.x == 0)
            for (int k = 0; k < SPLIT_NV; ++k) red[RED_OWN + k] = p[k];
    }

    // The window exchange (see the file comment).  On return RED_TOT holds
    // the system totals and, with_halo, the halo sites hold the spins their
    // owners had at the window end.
    __device__ void exchange(bool with_halo) {
        const uint32_t e = epoch++;
        const int par = (int)(e & 1u), R = a.split.R, G = a.split.G, B = a.split.bmax;
        const int t = (int)threadIdx.x;
        const uint32_t lim = a.spin_limit >> 4;
        // this system's published planes, parts one after another
        uint64_t *hx = a.split.hx + ((int64_t)par * R + sys) * (int64_t)G * 3 * B;
        uint64_t *px = a.split.px + ((int64_t)par * R + sys) * (SPLIT_MAX * 2 * SPLIT_NV);
        // RED_OWN written; every read of the spins in LDS done
        __syncthreads();
        if (t < 2 * SPLIT_NV) {
            const uint32_t *w = reinterpret_cast<const uint32_t *>(red + RED_OWN);
            ll_put(px + part * (2 * SPLIT_NV) + t, w[t], e);
        }
        if (with_halo) {
            uint64_t *ob = hx + (int64_t)part * 3 * B;
            for (int i = t; i < n_bnd; i += REG_THREADS) {
                const int li = lds_int(bnd_l + 4u * (uint32_t)i);
                const spin4<float> s = X.ld(base + 16u * (uint32_t)li);
                ll_put(ob + i, __float_as_uint(s.x), e);
                ll_put(ob + B + i, __float_as_uint(s.y), e);
                ll_put(ob + 2 * B + i, __float_as_uint(s.z), e);
            }
        }
        if (t < G * 2 * SPLIT_NV) {
            const uint64_t *w = px + t;  // part t / (2 SPLIT_NV), word t % (2 SPLIT_NV)
            reinterpret_cast<uint32_t *>(red + RED_PART)[t] = ll_wait(w, ll_load(w), e, lim, a.pt.err);
        }
        if (with_halo) {
            constexpr int U = 4;
            for (int i0 = t; i0 < n_halo; i0 += U * REG_THREADS) {
                int li[U];
                const uint64_t *w[U];
                uint64_t v[U][3];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t hp = halo_l + 8u * (uint32_t)min(i0 + u * REG_THREADS, n_halo - 1);
                    li[u] = lds_int(hp);
                    w[u] = hx + lds_int(hp + 4u);
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[u][k] = ll_load(w[u] + k * B);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (i0 + u * REG_THREADS < n_halo) {
                        spin4<float> s;
                        s.x = __uint_as_float(ll_wait(w[u], v[u][0], e, lim, a.pt.err));
                        s.y = __uint_as_float(ll_wait(w[u] + B, v[u][1], e, lim, a.pt.err));
                        s.z = __uint_as_float(ll_wait(w[u] + 2 * B, v[u][2], e, lim, a.pt.err));
                        s.w = 0.0f;
                        X.st(base + 16u * (uint32_t)li[u], s);
                    }
                }
            }
        }
        __syncthreads();
        if (t == 0)
            for (int k = 0; k < SPLIT_NV; ++k) {
                double s = red[RED_PART + k];
                for (int g = 1; g < G; ++g) s += red[RED_PART + g * SPLIT_NV + k];
                red[RED_TOT + k] = s;
            }
        __syncthreads();
    }
};

template <int K2C, int SPT, bool DIAG>
__global__ __launch_bounds__(REG_THREADS) void k_wg_split(WGArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *red = reinterpret_cast<double *>(smem);

    const int R = a.split.R;
    const int sys = (int)blockIdx.x % R, part = (int)blockIdx.x / R;
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    const int4 *runs = a.P.runs + ld.run_off;
    const LatView<float> L = lat_view<float>(ld, a.P.i32, runs, a.P.coefF);
    const LatView<double> Ld = lat_view<double>(ld, a.P.i32, runs, a.P.coefD);
    const SpinsG<float> G = spins_at<float>(a.spins, sd.spin_off);
    const int32_t *tab = a.split.tab;
    const int32_t *hdr = tab + part * SPLIT_HDR;
    const int n_loc = ldc(hdr + 0);
    const int32_t *locdev = tab + ldc(hdr + 10);

    const uint32_t base = lds_address(smem + RED_BYTES);
    SpinsV4 X;
    // stage the local sites and the zero pad site n_loc, then the boundary
    // list and the halo pairs
    wg_stage<8>(n_loc + 1, [&](int i) { return i < n_loc ? G.ld(locdev[i]) : spin4<float>{0, 0, 0, 0}; },
                [&](int i, const spin4<float> &v) { X.st(base + 16u * (uint32_t)i, v); });
    const int n_bnd = ldc(hdr + 3), n_halo = ldc(hdr + 4);
    const uint32_t bnd_l = base + 16u * (uint32_t)(n_loc + 1);
    const uint32_t halo_l = (bnd_l + 4u * (uint32_t)n_bnd + 7u) & ~7u;
    const int32_t *bnd = tab + ldc(hdr + 13), *halo = tab + ldc(hdr + 14);
    for (int i = threadIdx.x; i < n_bnd; i += REG_THREADS) lds_int_st(bnd_l + 4u * (uint32_t)i, bnd[i]);
    for (int i = threadIdx.x; i < 2 * n_halo; i += REG_THREADS) lds_int_st(halo_l + 4u * (uint32_t)i, halo[i]);
    SweeperS<K2C, SPT, DIAG> sw(L, Ld, a, base, red, sys, part, tab, hdr, bnd_l, halo_l);
    __syncthreads();
    wg_drive<WG_PT_CHAIN>(a, sys, sw);
    // every part read the start epoch before the first exchange, and the
    // leader gets here only after every part's last exchange
    if (threadIdx.x == 0 && part == 0) a.split.epoch[sys] = sw.epoch;
    // own sites back to the system's spins
    const int32_t *colstart = tab + ldc(hdr + 5), *owncnt = tab + ldc(hdr + 6);
    for (int c = 0; c < L.ncol; ++c) {
        const int cs = ldc(colstart + c), n = ldc(owncnt + c);
        for (int i = threadIdx.x; i < n; i += REG_THREADS)
            G.st(locdev[cs + i], X.ld(base + 16u * (uint32_t)(cs + i)));
    }
}

template <int K2C, int SPT, bool DIAG>
static hipError_t launch_split_t(int G, int R, size_t lds_bytes, const WGArgs &a, hipStream_t st) {
    const size_t bytes = RED_BYTES + lds_bytes;
    const void *fn = (const void *)k_wg_split<K2C, SPT, DIAG>;
    if (bytes > 65536) {
        hipError_t e = allow_large_lds(fn);
        if (e != hipSuccess) return e;
    }
    return launch_cooperative(fn, G * R, REG_THREADS, bytes, a, st);
}

template <int K2C>
static hipError_t launch_split_s(int spt, bool diag, int G, int R, size_t lds_bytes, const WGArgs &a,
                                 hipStream_t st) {
    if (spt == 4)
        return diag ? launch_split_t<K2C, 4, true>(G, R, lds_bytes, a, st)
                    : launch_split_t<K2C, 4, false>(G, R, lds_bytes, a, st);
    // 8 slots of 6 neighbours exceed the register budget of 1024 threads
    if constexpr (K2C <= 4) {
        if (spt == 8)
            return diag ? launch_split_t<K2C, 8, true>(G, R, lds_bytes, a, st)
                        : launch_split_t<K2C, 8, false>(G, R, lds_bytes, a, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_wg_split(int k2c, int spt, bool diag, int G, int R, size_t lds_bytes,
                           const WGArgs &a, hipStream_t st) {
    if (a.mode != WG_PT_CHAIN || G < 2 || G > SPLIT_MAX || a.split.G != G || a.split.R != R)
        return hipErrorInvalidValue;
    if (k2c == 3) return launch_split_s<3>(spt, diag, G, R, lds_bytes, a, st);
    if (k2c == 4) return launch_split_s<4>(spt, diag, G, R, lds_bytes, a, st);
    if (k2c == 6) return launch_split_s<6>(spt, diag, G, R, lds_bytes, a, st);
    return hipErrorInvalidValue;
}

}  // namespace csmc
