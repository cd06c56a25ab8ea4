// tile.hip -- one or several coloured sweeps of every system from LDS tiles,
// for periodic Bravais lattices whose neighbours and device order are
// arithmetic (the host plan and its conditions are in tile.cpp).
//
// A workgroup sweeps one tile of one system at a time:
//   1. stage the box (tile + halo) from the packed global records into LDS,
//      each site's device index computed from its cell and basis;
//   2. for each colour step of the launch's sweeps (nsw x ncol steps),
//      update every site of that colour in the step's region: the tile, plus
//      the halo sites that later steps of this tile read.  A partner's box
//      index is the site's index plus a per-(basis, slot) constant, so no
//      neighbour index is loaded.  The updates are the ones of the
//      colour-pass kernel (same functions, same Philox counter (reference
//      site, stream, sweep counter + the sweep's index in the launch,
//      METR)), so a halo site ends each step with the value its own tile
//      gives it;
//   3. store the tile's own sites.
// Every launch thus reads each record once (plus the halo) and writes it
// once for nsw sweeps, instead of once per colour pass with a gathered index
// per neighbour.  The launch reads one spin buffer and writes the other: a
// tile's halo must hold the values from before the launch, which a
// neighbouring tile of the same launch may already have replaced in place.
#include <type_traits>

#include "sweep.h"

namespace csmc {

// partner of slot k of the current basis: box index + delta, or the zero pad
// site for a slot that names no site.  Indices are in units of the box
// view's index (W per site: 3 words for packed FP32 records).
template <int W>
struct NbrD {
    static constexpr bool kKnowsK2 = true;
    const int32_t *d;  // delta row of the basis (wave-uniform)
    int K2, pad;
    __device__ __forceinline__ int operator()(int k, int i) const {
        if (k >= K2) return pad;
        const int v = ldc(d + k);
        // the scaled step stays a scalar operand of one vector add
        int vw = v * W;
        asm("" : "+s"(vw));
        return v == TILE_PAD ? pad : i + vw;
    }
    __device__ __forceinline__ int at(int k, int i) const { return (*this)(k, i); }
};

// a b for 0 <= a, b < 2^24 (a full-rate 24-bit multiply; the product must
// fit 32 bits)
// (written as the instructions: the compiler otherwise forms some of these
// products with the quarter-rate 32-bit multiplies)
__device__ __forceinline__ int mulu24(int a, int b) {
    int r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// a b + c for 0 <= a, b < 2^24
__device__ __forceinline__ int madu24(int a, int b, int c) {
    int r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// a / d for 0 <= a < 2^24 and 1 <= d < 2^24 from a float reciprocal rd of
// d: the float quotient is within one of the true one (exact for d = 1),
// and the remainder fixes it; the products fit 24-bit multiplies
__device__ __forceinline__ int fdiv(int a, int d, float rd) {
    int q = (int)((float)a * rd);
    const int r = a - mulu24(q, d);
    q += r < 0 ? -1 : (r >= d ? 1 : 0);
    return q;
}

__device__ __forceinline__ int imod(int a, int m) {
    const int r = a % m;
    return r < 0 ? r + m : r;
}

// byte offset of packed record d (3 x 4 or 3 x 8 bytes) as a shift and a
// shift-add (the compiler would otherwise multiply with a quarter-rate
// instruction)
template <class R>
__device__ __forceinline__ uint32_t rec_off(int d) {
    uint32_t q = (uint32_t)d << (sizeof(R) == 4 ? 2 : 3);
    asm("" : "+v"(q));
    return q + (q << 1);
}

// the packed record at byte offset off from a wave-uniform base: a 32-bit
// offset, so the address is base + offset without 64-bit arithmetic
template <class R>
__device__ __forceinline__ spin4<R> rec_ld(const unsigned char *base, uint32_t off) {
    const rec3<R> r = *reinterpret_cast<const rec3<R> *>(base + off);
    return spin4<R>{r.x, r.y, r.z, R(0)};
}
template <class R>
__device__ __forceinline__ void rec_st(unsigned char *base, uint32_t off, const spin4<R> &v) {
    *reinterpret_cast<rec3<R> *>(base + off) = rec3<R>{v.x, v.y, v.z};
}

// global cell of box-relative coordinate v along dimension d (one wrap)
__device__ __forceinline__ int wrap1(int v, int Ld) {
    v += v < 0 ? Ld : 0;
    return v - (v >= Ld ? Ld : 0);
}

// FP32 spins in LDS as packed {x, y, z} records indexed by word (3 per
// site): one address per site, the components at offsets 0, 4 and 8 (a
// stride of three words between lanes meets every bank once)
template <class R>
struct SpinsLds3 {
    R *p;
    __device__ __forceinline__ spin4<R> ld(int w) const {
        const R *q = p + w;
        return spin4<R>{q[0], q[1], q[2], R(0)};
    }
    __device__ __forceinline__ void st(int w, const spin4<R> &v) const {
        R *q = p + w;
        q[0] = v.x;
        q[1] = v.y;
        q[2] = v.z;
    }
};

// the LDS view of the box: packed records in FP32 (index = 3 x site), and
// component arrays in FP64 (index = site; 24-byte records would meet banks
// twice per wave)
template <class R>
constexpr int box_w() {
    return sizeof(R) == 4 ? 3 : 1;
}
template <class R>
using BoxView = std::conditional_t<sizeof(R) == 4, SpinsLds3<R>, SpinsSoA<R>>;
template <class R>
__device__ __forceinline__ BoxView<R> box_view(R *sx, int nrec) {
    if constexpr (sizeof(R) == 4)
        return SpinsLds3<R>{sx};
    else
        return SpinsSoA<R>{sx, sx + nrec, sx + 2 * nrec};
}

// remainder of 0 <= a < 2^24 by 1 <= m < 2^24 (rm: float reciprocal of m)
__device__ __forceinline__ int fmodu(int a, int m, float rm) { return a - mulu24(fdiv(a, m, rm), m); }

// A walk over the cells of an n0 x n1 x n2 block, row-major, by a fixed
// stride: the position (u0, u1, u2) of cell c, then of c + stride, ...
// (stride and every index below 2^24; one carry per dimension and step
// because the stride's own digits d1 < n1 and d2 < n2)
struct Walk3 {
    int u0, u1, u2, d0, d1, d2, n1, n2;
    __device__ __forceinline__ Walk3(int c, int stride, int n1_, int n2_) : n1(n1_), n2(n2_) {
        const int n12 = n1 * n2;
        const float r12 = 1.0f / (float)n12, r2 = 1.0f / (float)n2;
        u0 = fdiv(c, n12, r12);
        const int rc = c - mulu24(u0, n12);
        u1 = fdiv(rc, n2, r2);
        u2 = rc - mulu24(u1, n2);
        d0 = fdiv(stride, n12, r12);
        const int rs = stride - mulu24(d0, n12);
        d1 = fdiv(rs, n2, r2);
        d2 = rs - mulu24(d1, n2);
    }
    __device__ __forceinline__ void step() {
        u2 += d2;
        const int c2 = u2 >= n2 ? 1 : 0;
        u2 -= c2 ? n2 : 0;
        u1 += d1 + c2;
        const int c1 = u1 >= n1 ? 1 : 0;
        u1 -= c1 ? n1 : 0;
        u0 += d0 + c1;
    }
};

// KIND 0: overrelaxation, 1: deterministic, 2: Metropolis (a.kind picks the
// uniform or the cone proposal).  FLAT: the first of the three positions has
// extent 1 for every lattice of the launch (two-dimensional lattices), so its
// coordinate is 0 throughout.  Index products use 24-bit multiplies: the plan
// keeps every extent, the row count and the box below 2^24.
template <class R, bool SELF, int KIND, int NT, bool FLAT>
__global__ __launch_bounds__(NT) void k_tile(TileArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr uint32_t RB = 3 * sizeof(R);  // bytes per packed record
    const int sys = a.sys0 + (int)blockIdx.y;
    const SysDev sd = a.P.sys[sys];
    const TileDev t = a.td[sd.lat];
    if ((int)blockIdx.x >= t.ntiles) return;
    const LatDev ld = a.P.lats[sd.lat];
    const LatView<R> L = lat_view<R>(ld, a.P.i32, a.P.runs + ld.run_off, coef_of<R>(a.P));
    const unsigned char *gin = reinterpret_cast<const unsigned char *>(a.spins_in) + sd.spin_off * RB;
    unsigned char *gout = reinterpret_cast<unsigned char *>(a.spins_out) + sd.spin_off * RB;
    const int nbox = t.nbox, nb = t.nb;
    R *sx = reinterpret_cast<R *>(smem);
    const BoxView<R> X = box_view<R>(sx, nbox + 1);
    // run bases in LDS after the spins: runbase[r * nb + b]
    int *sbase = reinterpret_cast<int *>(sx + 3 * (nbox + 1));
    const int32_t *eb = a.tab + t.tab_off, *runbase = eb + nb, *delta = runbase + t.p * nb;
    const int32_t *rect = delta + nb * t.K2, *rawc = rect + t.nsw * t.ncol * nb * 6;
    for (int k = threadIdx.x; k < t.p * nb; k += NT) sbase[k] = runbase[k];
    constexpr int W = box_w<R>();  // view index units per site
    if (threadIdx.x == 0) X.st(nbox * W, spin4<R>{R(0), R(0), R(0), R(0)});
    const float r_p = 1.0f / (float)t.p;
    const int l2p = t.cz ? 0 : t.L[2] / t.p;  // cells of one raw colour per row
    int acc = 0;
    __syncthreads();  // sbase
    for (int tile = blockIdx.x; tile < t.ntiles; tile += gridDim.x) {
        int g0[3];  // global cell of box cell 0 (before wrapping)
        {
            int r = tile;
            for (int d = 2; d >= 0; --d) {
                const int q = r / t.nt[d];
                g0[d] = (r - q * t.nt[d]) * t.T[d] + t.lo[d];
                r = q;
            }
        }
        // device index of basis b (wave-uniform) at global cell x (cz: base =
        // the run base of the basis, read once per basis plane)
        auto dev = [&](int b, int base, int x0, int x1, int x2) {
            const int row = FLAT ? x1 : madu24(x0, t.L[1], x1);
            if (t.cz) return base + madu24(row, t.L[2], x2);
            const int r = fmodu(ldc(eb + b) + (FLAT ? 0 : mulu24(t.c[0], x0)) + mulu24(t.c[1], x1) +
                                    mulu24(t.c[2], x2),
                                t.p, r_p);
            return sbase[r * nb + b] + mulu24(row, l2p) + fdiv(x2, t.p, r_p);
        };
        // 1. the box, one basis plane at a time, four loads in flight (the
        // plane's last cells one at a time)
        for (int b = 0; b < nb; ++b) {
            const int bo = b * t.boxcells;
            const int base = t.cz ? sbase[ldc(eb + b) * nb + b] : 0;
            Walk3 wk((int)threadIdx.x, NT, t.B[1], t.B[2]);
            // the record of the walk's cell, then the walk's next cell
            auto load = [&]() {
                const int x0 = FLAT ? 0 : wrap1(g0[0] + wk.u0, t.L[0]);
                const int d = dev(b, base, x0, wrap1(g0[1] + wk.u1, t.L[1]), wrap1(g0[2] + wk.u2, t.L[2]));
                const spin4<R> v = rec_ld<R>(gin, rec_off<R>(d));
                wk.step();
                return v;
            };
            int c0 = threadIdx.x;
            for (; c0 + 3 * NT < t.boxcells; c0 += 4 * NT) {
                spin4<R> v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] = load();
#pragma unroll
                for (int q = 0; q < 4; ++q) X.st((bo + c0 + q * NT) * W, v[q]);
            }
            for (; c0 < t.boxcells; c0 += NT) X.st((bo + c0) * W, load());
        }
        __syncthreads();
        // 2. the colour steps
        const bool metro = KIND == 2;
        const R T = metro ? (R)a.T_sys[sys] : R(1);
        const R invT = R(1) / T;
        const R sigma = (metro && a.sigma_sys) ? (R)a.sigma_sys[sys] : R(60);
        const uint32_t ctr = metro ? (a.ctr ? a.ctr[sys] : a.ctr0) : 0u;
        // step st: sweep st / ncol of the launch, colour st % ncol
        for (int st = 0, col = 0, sw = 0; st < t.nsw * t.ncol; ++st, col = (col + 1 == t.ncol ? 0 : col + 1),
                 sw += (col == 0 ? 1 : 0)) {
            const int r0 = ldc(rawc + col);
            for (int b = 0; b < nb; ++b) {
                const int32_t *rc = rect + (st * nb + b) * 6;
                const int lo0 = ldc(rc), hi0 = ldc(rc + 1), lo1 = ldc(rc + 2), hi1 = ldc(rc + 3);
                const int lo2 = ldc(rc + 4), hi2 = ldc(rc + 5);
                if (lo0 >= hi0 || lo1 >= hi1 || lo2 >= hi2) continue;
                const NbrD<W> nbr{delta + b * t.K2, t.K2, nbox * W};
                // cells per row: all of them (colour by basis), or every p-th
                const int w2 = t.cz ? hi2 - lo2 : (hi2 - lo2 + t.p - 1) / t.p;
                const int n1 = hi1 - lo1;
                const int cnt = (hi0 - lo0) * n1 * w2;
                const int e0 = ldc(eb + b);
                const int bo = b * t.boxcells;
                Walk3 wk((int)threadIdx.x, NT, n1, w2);
                for (int w = threadIdx.x; w < cnt; w += NT, wk.step()) {
                    const int u0 = FLAT ? 0 : lo0 + wk.u0, u1 = lo1 + wk.u1;
                    int u2 = lo2 + wk.u2;
                    if (!t.cz) {
                        // the cells of raw colour r0 in this row: c2 (lo2b + u2)
                        // = r0 - e_b - c0 (lo0b + u0) - c1 (lo1b + u1) mod p
                        // (tile origins are multiples of the pattern's period)
                        const int s = r0 - e0 - (FLAT ? 0 : t.c[0] * (t.lo[0] + u0)) - t.c[1] * (t.lo[1] + u1);
                        u2 = lo2 + imod(s * t.cinv - t.lo[2] - lo2, t.p) + wk.u2 * t.p;
                        if (u2 >= hi2) continue;
                    }
                    const int cell = FLAT ? madu24(u1, t.B[2], u2) : madu24(madu24(u0, t.B[1], u1), t.B[2], u2);
                    const int li = mulu24(bo + cell, W);
                    if constexpr (KIND == 0) {
                        or_update<R, false, 0>(L, nbr, X, b, li);
                    } else if constexpr (KIND == 1) {
                        det_update<R, false, 0>(L, nbr, X, b, li);
                    } else {
                        const int x0 = FLAT ? 0 : wrap1(g0[0] + u0, t.L[0]), x1 = wrap1(g0[1] + u1, t.L[1]);
                        const int x2 = wrap1(g0[2] + u2, t.L[2]);
                        const bool own = (FLAT || (unsigned)(u0 + t.lo[0]) < (unsigned)t.T[0]) &&
                                         (unsigned)(u1 + t.lo[1]) < (unsigned)t.T[1] &&
                                         (unsigned)(u2 + t.lo[2]) < (unsigned)t.T[2];
                        const uint32_t ref =
                            (uint32_t)madu24(madu24(b * t.L[0] + x0, t.L[1], x1), t.L[2], x2);
                        const spin4<R> s = X.ld(li);
                        const uint4 r = philox4x32_10(make_uint4(ref, (uint32_t)sd.stream, ctr + (uint32_t)sw,
                                                                 TAG_METRO), a.k0,
                                                      a.k1);
                        const spin4<R> n = metro_propose<R>(L, s, r, a.kind, sigma);
                        const R dE = metro_eval<R, SELF, false, 0>(L, nbr, X, b, li, s, n);
                        const int ok = metro_accept<R>(dE, r, T, invT);
                        metro_commit<R>(L, X, li, ok, s, n);
                        acc += own ? ok : 0;
                    }
                }
            }
            __syncthreads();
        }
        // 3. the tile's own sites, one basis plane at a time
        for (int b = 0; b < nb; ++b) {
            const int base = t.cz ? sbase[ldc(eb + b) * nb + b] : 0;
            Walk3 wk((int)threadIdx.x, NT, t.T[1], t.T[2]);
            for (int c = threadIdx.x; c < t.tcells; c += NT, wk.step()) {
                const int u0 = FLAT ? 0 : wk.u0 - t.lo[0], u1 = wk.u1 - t.lo[1], u2 = wk.u2 - t.lo[2];
                const int li = b * t.boxcells +
                               (FLAT ? madu24(u1, t.B[2], u2) : madu24(madu24(u0, t.B[1], u1), t.B[2], u2));
                // g0 + u is inside the lattice for the tile's own cells
                const int d = dev(b, base, FLAT ? 0 : g0[0] + u0, g0[1] + u1, g0[2] + u2);
                rec_st<R>(gout, rec_off<R>(d), X.ld(mulu24(li, W)));
            }
        }
        __syncthreads();  // the next tile's box overwrites LDS
    }
    if constexpr (KIND == 2) {
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0 && acc) atomicAdd(&a.acc[sys], (uint32_t)acc);
    }
}

template <class R, bool SELF, int KIND, int NT, bool FLAT>
static hipError_t launch_tile_n(dim3 grid, size_t lds, const TileArgs &a, hipStream_t st) {
    if (lds > 65536) {
        const hipError_t e = allow_large_lds((const void *)k_tile<R, SELF, KIND, NT, FLAT>);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_tile<R, SELF, KIND, NT, FLAT>), grid, dim3(NT), lds, st, a);
    return hipGetLastError();
}

// 256 threads a tile: 512 and 1024 measured 17 % and 60 % slower on the
// honeycomb L=512 probe (more of each colour step's last pass idle)
template <class R, bool SELF, int KIND>
static hipError_t launch_tile_t(dim3 grid, size_t lds, const TileArgs &a, hipStream_t st) {
    if (a.flat) return launch_tile_n<R, SELF, KIND, 256, true>(grid, lds, a, st);
    return launch_tile_n<R, SELF, KIND, 256, false>(grid, lds, a, st);
}

// The field uses the slot loop of the colour-pass kernel (slots in groups of
// four, each matrix row's sum added to the field): the unrolled slot-count
// forms add the products in another order, and the tiled sweep must equal
// the colour-pass one bit for bit.  Overrelaxation and deterministic sweeps
// never name the site itself in a proposal, so SELF matters to Metropolis
// only.
template <class R>
static hipError_t launch_tile_r(bool self, int kind, dim3 grid, size_t lds, const TileArgs &a, hipStream_t st) {
    if (kind == CSMC_OVERRELAX) return launch_tile_t<R, false, 0>(grid, lds, a, st);
    if (kind == CSMC_DETERMINISTIC) return launch_tile_t<R, false, 1>(grid, lds, a, st);
    if (self) return launch_tile_t<R, true, 2>(grid, lds, a, st);
    return launch_tile_t<R, false, 2>(grid, lds, a, st);
}

hipError_t launch_tile(int prec, bool self, int nsys, int grid_x, size_t lds_bytes, const TileArgs &a,
                       hipStream_t st) {
    const dim3 grid(grid_x, nsys);
    if (prec == 32) return launch_tile_r<float>(self, a.kind, grid, lds_bytes, a, st);
    return launch_tile_r<double>(self, a.kind, grid, lds_bytes, a, st);
}

}  // namespace csmc
