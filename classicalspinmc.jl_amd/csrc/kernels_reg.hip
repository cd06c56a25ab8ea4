// kernels_reg.hip -- register-table storage mode of the persistent
// workgroup-per-system sweep kernel (FP32 state, bilinear-only lattices).
//
// One workgroup of REG_THREADS threads owns one system.  Its spins sit in
// LDS as float4 records (16 B per site, plus the zero pad site), and each
// thread keeps, in registers, the LDS byte offsets of the sites it updates
// and of their K2C bilinear neighbours: thread t owns site first + t of every
// slot (a REG_THREADS-site piece of one run of one colour class).  A site
// update is then one 16-byte LDS read per spin involved and one 16-byte
// write, with no neighbour-table traffic and no address arithmetic in the
// sweep.  The schedule (wg_drive) and the per-site physics (device.h) are the
// ones every other storage mode uses.
#include "reg.h"

namespace csmc {

// SPC > 0: the lattice has two colour classes of SPC full REG_THREADS-site
// slots each (SPT = 2 SPC), so every thread has a site in every slot and no
// update needs an ownership mask; SPC = 0: any layout.
template <int K2C, int SPT, bool DIAG, int SPC>
struct SweeperR {
    const LatView<float> &L;
    const LatView<double> &Ld;
    SpinsV4 X;
    double *red;
    const int32_t *slots;    // int4 records {first site, run end, basis, colour}
    const int32_t *colslot;  // first slot of each colour [ncol + 1], then the
                             // basis shared by a colour's slots (-1 if mixed) [ncol]
    int nslots;
    uint32_t stream, k0, k1;
    uint32_t base;           // LDS address of site 0
    uint32_t valid;          // bit q: this thread owns a site in slot q
    uint32_t own[SPT];       // LDS address of the owned site (the zero pad site if none)
    uint32_t nb[SPT][K2C];   // LDS address of each bilinear neighbour (pad site if absent)

    // spins a site update reads: its own and its K2C neighbours
    struct Site {
        spin4<float> s;
        spin4<float> sj[K2C];
    };

    __device__ SweeperR(const LatView<float> &L_, const LatView<double> &Ld_, uint32_t base_,
                        double *red_, const int32_t *slots_, const int32_t *colslot_, int nslots_,
                        uint32_t stream_, uint32_t k0_, uint32_t k1_)
        : L(L_), Ld(Ld_), X(), red(red_), slots(slots_), colslot(colslot_), nslots(nslots_),
          stream(stream_), k0(k0_), k1(k1_), base(base_), valid(0) {
        // every table load is unconditional (slot index and site clamped to
        // valid entries) so all of them are in flight together
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int qq = q < nslots ? q : nslots - 1;
            const int first = ldc(slots + 4 * qq), end = ldc(slots + 4 * qq + 1);
            const int i = first + (int)threadIdx.x;
            const bool ok = q < nslots && i < end;
            const int ic = ok ? i : first;
            valid |= ok ? (1u << q) : 0u;
            own[q] = pin_address(base + 16u * (uint32_t)(ok ? i : L.N));
#pragma unroll
            for (int k = 0; k < K2C; ++k)
                nb[q][k] = pin_address(base + 16u * (uint32_t)L.nbr2[(int64_t)k * L.N + ic]);
        }
    }

    __device__ int n_sites() const { return L.N; }
    __device__ bool leader() const { return true; }  // the only workgroup of its system
    __device__ bool owns(int q) const { return (valid >> q) & 1u; }
    // a sweep needs no ownership test when every slot is full
    __device__ bool sweeps(int q) const { return SPC > 0 || owns(q); }
    __device__ int basis(int q) const { return ldc(slots + 4 * q + 2); }
    __device__ int site(int q) const { return (int)((own[q] - base) >> 4); }

    __device__ __forceinline__ Site load(int q) const {
        Site v;
        v.s = X.ld(own[q]);
#pragma unroll
        for (int k = 0; k < K2C; ++k) v.sj[k] = X.ld(nb[q][k]);
        return v;
    }

    // one coloured sweep: colour classes in turn, a workgroup barrier after
    // each.  Every colour class has a single basis (the host only picks this
    // mode then), so its couplings are loaded once for the colour.  Slots are
    // taken two at a time, both slots' spins loaded before either is updated
    // (the sites are independent), so two LDS round trips overlap;
    // upd(q, basis, site, M) updates slot q with couplings M.
    template <class UPD>
    __device__ __forceinline__ void colour_sweep(UPD &&upd) {
        for (int c = 0; c < L.ncol; ++c) {
            const int qb = ldc(colslot + c), qe = ldc(colslot + c + 1);
            const int cb = ldc(colslot + L.ncol + 1 + c);
            BilCoef<float, K2C, DIAG> M;
            M.load(L, cb);
            slot_pairs(qb, qe, [&](int q, const Site &v) { upd(q, cb, v, M); });
            __syncthreads();
        }
    }

    template <class F>
    __device__ __forceinline__ void slot_pairs(int qb, int qe, F &&f) {
#pragma unroll
        for (int q = 0; q < SPT; q += 2) {
            const bool in0 = q >= qb && q < qe;
            const bool in1 = q + 1 < SPT && q + 1 >= qb && q + 1 < qe;
            if (in0 && in1) {
                const Site v0 = load(q), v1 = load(q + 1);
                if (sweeps(q)) f(q, v0);
                if (sweeps(q + 1)) f(q + 1, v1);
            } else if (in0) {
                if (sweeps(q)) f(q, load(q));
            } else if (in1) {
                if (sweeps(q + 1)) f(q + 1, load(q + 1));
            }
        }
    }

    __device__ void or_sweep() {
        colour_sweep([&](int q, int b, const Site &v, const BilCoef<float, K2C, DIAG> &M) {
            spin4<float> n;
            float F[3], H[3];
            M.field(v.sj, F);
            local_field<float, float>(L, b, v.s, F, H);
            // or_reflect, with the H == 0 case (no change) taken by a select
            // instead of a branch: the spin is stored back unchanged
            const bool zero = H[0] == 0.0f && H[1] == 0.0f && H[2] == 0.0f;
            const float hh = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
            const float proj = Math<float>::div(2.0f * (v.s.x * H[0] + v.s.y * H[1] + v.s.z * H[2]), hh);
            n.x = zero ? v.s.x : -v.s.x + proj * H[0];
            n.y = zero ? v.s.y : -v.s.y + proj * H[1];
            n.z = zero ? v.s.z : -v.s.z + proj * H[2];
            n.w = 0.0f;
            X.st(own[q], n);
        });
    }

    __device__ void det_sweep() {
        colour_sweep([&](int q, int b, const Site &v, const BilCoef<float, K2C, DIAG> &M) {
            spin4<float> n;
            float F[3], H[3];
            M.field(v.sj, F);
            local_field<float, float>(L, b, v.s, F, H);
            if (det_align<float>(L, H, n)) X.st(own[q], n);
        });
    }

    __device__ double metro(double T, int kind, double sigma, uint32_t ctr) {
        double w[1] = {(double)metro_local(T, kind, sigma, ctr)};
        block_sum<1>(w, red);
        return w[0];
    }

    // one Metropolis sweep; returns this thread's accepted count
    __device__ int metro_local(double T, int kind, double sigma, uint32_t ctr) {
        const float Tf = (float)T, invT = 1.0f / Tf, sg = (float)sigma;
        int acc = 0;
        colour_sweep([&](int q, int b, const Site &v, const BilCoef<float, K2C, DIAG> &M) {
            // Philox counter: the reference index of the site
            const uint32_t ref = (uint32_t)L.ref[site(q)];
            uint4 r = philox4x32_10(make_uint4(ref, stream, ctr, TAG_METRO), k0, k1);
            spin4<float> n = metro_propose<float>(L, v.s, r, kind, sg);
            float F[3];
            M.field(v.sj, F);
            int a = metro_accept<float>(metro_dE<float>(L, b, v.s, n, F), r, Tf, invT);
            metro_commit<float>(L, X, own[q], a, v.s, n);
            acc += a;
        });
        return acc;
    }

    // this thread's FP64 energy (bilinear / 2 + Zeeman + onsite: the lattice
    // has no cubic or quartic terms) and summed spin
    __device__ void thread_energy(double &e, double m[3]) {
        double bil = 0.0, zee = 0.0, ons = 0.0;
        m[0] = m[1] = m[2] = 0.0;
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            if (owns(q)) {
                const int b = basis(q);
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
                m[0] += sx;
                m[1] += sy;
                m[2] += sz;
            }
        }
        e = bil / 2 + zee + ons;
    }

    __device__ double metro_energy(double T, int kind, double sigma, uint32_t ctr, bool with_m,
                                   double &E, double m[3]) {
        const int acc = metro_local(T, kind, sigma, ctr);
        double e, mm[3];
        thread_energy(e, mm);
        if (with_m) {
            double v[5] = {(double)acc, e, mm[0], mm[1], mm[2]};
            block_sum<5>(v, red);
            E = v[1];
            m[0] = v[2];
            m[1] = v[3];
            m[2] = v[4];
            return v[0];
        }
        double v[2] = {(double)acc, e};
        block_sum<2>(v, red);
        E = v[1];
        return v[0];
    }

    // the summed spin over the same per-thread sites, in the same order, as
    // thread_energy, so both give the same FP64 sums
    __device__ void magnet(double out[3]) {
        double m[3] = {0, 0, 0};
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            if (owns(q)) {
                const spin4<float> s = X.ld(own[q]);
                m[0] += s.x;
                m[1] += s.y;
                m[2] += s.z;
            }
        }
        block_sum<3>(m, red);
        for (int k = 0; k < 3; ++k) out[k] = m[k];
    }
};

template <int K2C, int SPT, bool DIAG, int MODE, int SPC>
__global__ __launch_bounds__(REG_THREADS) void k_wg_reg(WGArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *red = reinterpret_cast<double *>(smem);  // reduction scratch

    const int sys = blockIdx.x;
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    const int4 *runs = a.P.runs + ld.run_off;
    const LatView<float> L = lat_view<float>(ld, a.P.i32, runs, a.P.coefF);
    const LatView<double> Ld = lat_view<double>(ld, a.P.i32, runs, a.P.coefD);
    const SpinsG<float> G = spins_at<float>(a.spins, sd.spin_off);
    const int N = L.N;

    const uint32_t base = lds_address(smem + RED_BYTES);
    SpinsV4 X;
    // stage the system and its zero pad site N into LDS
    wg_stage<8>(N + 1, [&](int i) { return G.ld(i); },
                [&](int i, const spin4<float> &v) { X.st(base + 16u * (uint32_t)i, v); });
    SweeperR<K2C, SPT, DIAG, SPC> sw(L, Ld, base, red,
                                reinterpret_cast<const int32_t *>(a.P.runs + ld.slot_off),
                                a.P.i32 + ld.colslot_off, ld.nslots, (uint32_t)sd.stream, a.k0,
                                a.k1);
    __syncthreads();
    wg_drive<MODE>(a, sys, sw);
    // the exchange goes first so that its fence writes back only the small
    // PT state, not the spins stored below
    if (MODE == WG_PT && a.exchange) pt_exchange_last(a, reinterpret_cast<int *>(red + RED_FLAG));
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += REG_THREADS) G.st(i, X.ld(base + 16u * (uint32_t)i));
}

template <int K2C, int SPT, bool DIAG, int MODE, int SPC>
static hipError_t launch_reg_t(int nsys, size_t lds_bytes, const WGArgs &a, hipStream_t st) {
    const size_t bytes = RED_BYTES + lds_bytes;
    const void *fn = (const void *)k_wg_reg<K2C, SPT, DIAG, MODE, SPC>;
    if (bytes > 65536) {
        hipError_t e = allow_large_lds(fn);
        if (e != hipSuccess) return e;
    }
    if constexpr (MODE == WG_PT_CHAIN) return launch_cooperative(fn, nsys, REG_THREADS, bytes, a, st);
    hipLaunchKernelGGL((k_wg_reg<K2C, SPT, DIAG, MODE, SPC>), dim3(nsys), dim3(REG_THREADS), bytes, st, a);
    return hipGetLastError();
}

template <int K2C, int SPT, bool DIAG, int SPC>
static hipError_t launch_reg_m(int nsys, size_t lds_bytes, const WGArgs &a, hipStream_t st) {
    if (a.mode == WG_PT) return launch_reg_t<K2C, SPT, DIAG, WG_PT, SPC>(nsys, lds_bytes, a, st);
    if (a.mode == WG_PT_CHAIN)
        return launch_reg_t<K2C, SPT, DIAG, WG_PT_CHAIN, SPC>(nsys, lds_bytes, a, st);
    if (a.mode == WG_ANNEAL) return launch_reg_t<K2C, SPT, DIAG, WG_ANNEAL, SPC>(nsys, lds_bytes, a, st);
    return launch_reg_t<K2C, SPT, DIAG, WG_SWEEPS, SPC>(nsys, lds_bytes, a, st);
}

template <int K2C, int SPT>
static hipError_t launch_reg_d(bool diag, bool two_colour, int nsys, size_t lds_bytes,
                               const WGArgs &a, hipStream_t st) {
    // the fixed two-colour layout is compiled for diagonal couplings with 3
    // or 4 neighbours (honeycomb, square and their Kitaev / XXZ variants)
    if constexpr (K2C <= 4) {
        if (diag && two_colour) return launch_reg_m<K2C, SPT, true, SPT / 2>(nsys, lds_bytes, a, st);
    }
    return diag ? launch_reg_m<K2C, SPT, true, 0>(nsys, lds_bytes, a, st)
                : launch_reg_m<K2C, SPT, false, 0>(nsys, lds_bytes, a, st);
}

template <int K2C>
static hipError_t launch_reg_s(int spt, bool diag, bool two_colour, int nsys, size_t lds_bytes,
                               const WGArgs &a, hipStream_t st) {
    if (spt == 4) return launch_reg_d<K2C, 4>(diag, two_colour, nsys, lds_bytes, a, st);
    // 8 slots of 6 neighbours exceed the register budget of 1024 threads
    if constexpr (K2C <= 4) {
        if (spt == 8) return launch_reg_d<K2C, 8>(diag, two_colour, nsys, lds_bytes, a, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_wg_reg(int k2c, int spt, bool diag, bool two_colour, int nsys, size_t lds_bytes,
                         const WGArgs &a, hipStream_t st) {
    if (k2c == 3) return launch_reg_s<3>(spt, diag, two_colour, nsys, lds_bytes, a, st);
    if (k2c == 4) return launch_reg_s<4>(spt, diag, two_colour, nsys, lds_bytes, a, st);
    if (k2c == 6) return launch_reg_s<6>(spt, diag, two_colour, nsys, lds_bytes, a, st);
    return hipErrorInvalidValue;
}

}  // namespace csmc
