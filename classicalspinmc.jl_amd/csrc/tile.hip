// tile.hip -- one coloured sweep of every system from LDS tiles, for periodic
// Bravais lattices whose neighbours and device order are arithmetic (the
// host plan and its conditions are in tile.cpp).
//
// A workgroup sweeps one tile of one system at a time:
//   1. stage the box (tile + halo) from the packed global records into LDS,
//      each site's device index computed from its cell and basis;
//   2. for each colour step, update every site of that colour in the step's
//      region: the tile, plus the halo sites that later steps of this tile
//      read.  A partner's box index is the site's index plus a per-(basis,
//      slot) constant, so no neighbour index is loaded.  The updates are the
//      ones of the colour-pass kernel (same functions, same Philox counter
//      (reference site, stream, sweep counter, METR)), so a halo site ends
//      the step with the value its own tile gives it;
//   3. store the tile's own sites.
// Every system's sweep thus reads each record once (plus the halo) and writes
// it once, instead of once per colour pass with a gathered index per
// neighbour.  The sweep reads one spin buffer and writes the other: a tile's
// halo must hold the values from before the sweep, which a neighbouring tile
// of the same launch may already have replaced in place.
#include "sweep.h"

namespace csmc {

// partner of slot k of the current basis: box index + delta, or the zero pad
// site for a slot that names no site
struct NbrD {
    const int32_t *d;  // delta row of the basis (wave-uniform)
    int K2, pad;
    __device__ __forceinline__ int operator()(int k, int i) const {
        if (k >= K2) return pad;
        const int v = ldc(d + k);
        return v == TILE_PAD ? pad : i + v;
    }
    __device__ __forceinline__ int at(int k, int i) const { return (*this)(k, i); }
};

// a / d for 0 <= a < 2^24 and 1 <= d < 2^24 from a float reciprocal rd of
// d: the float quotient is within one of the true one (exact for d = 1),
// and the remainder fixes it; the products fit 24-bit multiplies
__device__ __forceinline__ int fdiv(int a, int d, float rd) {
    int q = (int)((float)a * rd);
    const int r = a - __mul24(q, d);
    q += r < 0 ? -1 : (r >= d ? 1 : 0);
    return q;
}

__device__ __forceinline__ int imod(int a, int m) {
    const int r = a % m;
    return r < 0 ? r + m : r;
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

// KIND 0: overrelaxation, 1: deterministic, 2: Metropolis (a.kind picks the
// uniform or the cone proposal)
template <class R, bool SELF, int KIND>
__global__ __launch_bounds__(256) void k_tile(TileArgs a) {
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
    const SpinsSoA<R> X{sx, sx + (nbox + 1), sx + 2 * (nbox + 1)};
    // run bases in LDS after the spins: runbase[r * nb + b]
    int *sbase = reinterpret_cast<int *>(sx + 3 * (nbox + 1));
    const int32_t *eb = a.tab + t.tab_off, *runbase = eb + nb, *delta = runbase + t.p * nb;
    const int32_t *rect = delta + nb * t.K2, *rawc = rect + t.nsw * t.ncol * nb * 6;
    for (int k = threadIdx.x; k < t.p * nb; k += blockDim.x) sbase[k] = runbase[k];
    if (threadIdx.x == 0) X.st(nbox, spin4<R>{R(0), R(0), R(0), R(0)});
    const float r_B1 = 1.0f / (float)t.B[1], r_B2 = 1.0f / (float)t.B[2];
    const float r_T1 = 1.0f / (float)t.T[1], r_T2 = 1.0f / (float)t.T[2];
    const int cells_row = t.B[1] * t.B[2], tcells = t.tcells, trow = t.T[1] * t.T[2];
    const float r_cr = 1.0f / (float)cells_row, r_tr = 1.0f / (float)trow;
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
        // device index of basis b at global cell x
        auto dev = [&](int b, int x0, int x1, int x2) {
            const int ci = (x0 * t.L[1] + x1) * t.L[2] + x2;
            if (t.cz) return sbase[ldc(eb + b) * nb + b] + ci;
            const int r = imod(ldc(eb + b) + t.c[0] * x0 + t.c[1] * x1 + t.c[2] * x2, t.p);
            return sbase[r * nb + b] + (ci / t.L[2]) * (t.L[2] / t.p) + x2 / t.p;
        };
        // 1. the box, one basis plane at a time, four loads in flight
        for (int b = 0; b < nb; ++b) {
            const int bo = b * t.boxcells;
            for (int c0 = threadIdx.x; c0 < t.boxcells; c0 += 4 * blockDim.x) {
                spin4<R> v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int c = min(c0 + q * (int)blockDim.x, t.boxcells - 1);
                    const int u0 = fdiv(c, cells_row, r_cr), rc = c - u0 * cells_row;
                    const int u1 = fdiv(rc, t.B[2], r_B2), u2 = rc - u1 * t.B[2];
                    const int d = dev(b, wrap1(g0[0] + u0, t.L[0]), wrap1(g0[1] + u1, t.L[1]),
                                      wrap1(g0[2] + u2, t.L[2]));
                    v[q] = rec_ld<R>(gin, (uint32_t)d * RB);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (c0 + q * (int)blockDim.x < t.boxcells) X.st(bo + c0 + q * (int)blockDim.x, v[q]);
            }
        }
        (void)r_B1;
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
                const NbrD nbr{delta + b * t.K2, t.K2, nbox};
                // cells per row: all of them (colour by basis), or every p-th
                const int w2 = t.cz ? hi2 - lo2 : (hi2 - lo2 + t.p - 1) / t.p;
                const int n1 = hi1 - lo1;
                const int cnt = (hi0 - lo0) * n1 * w2;
                const int e0 = ldc(eb + b);
                const float r_w2 = 1.0f / (float)w2, r_n1 = 1.0f / (float)n1;
                for (int w = threadIdx.x; w < cnt; w += blockDim.x) {
                    const int row = fdiv(w, w2, r_w2), j = w - row * w2;
                    const int q1 = fdiv(row, n1, r_n1);
                    const int u0 = lo0 + q1, u1 = lo1 + (row - q1 * n1);
                    int u2 = lo2 + j;
                    if (!t.cz) {
                        // the cells of raw colour r0 in this row: c2 (lo2b + u2)
                        // = r0 - e_b - c0 (lo0b + u0) - c1 (lo1b + u1) mod p
                        // (tile origins are multiples of the pattern's period)
                        const int s = r0 - e0 - t.c[0] * (t.lo[0] + u0) - t.c[1] * (t.lo[1] + u1);
                        u2 = lo2 + imod(s * t.cinv - t.lo[2] - lo2, t.p) + j * t.p;
                        if (u2 >= hi2) continue;
                    }
                    const int li = b * t.boxcells + (u0 * t.B[1] + u1) * t.B[2] + u2;
                    if constexpr (KIND == 0) {
                        or_update<R, false, 0>(L, nbr, X, b, li);
                    } else if constexpr (KIND == 1) {
                        det_update<R, false, 0>(L, nbr, X, b, li);
                    } else {
                        const int x0 = wrap1(g0[0] + u0, t.L[0]), x1 = wrap1(g0[1] + u1, t.L[1]);
                        const int x2 = wrap1(g0[2] + u2, t.L[2]);
                        const bool own = (unsigned)(u0 + t.lo[0]) < (unsigned)t.T[0] &&
                                         (unsigned)(u1 + t.lo[1]) < (unsigned)t.T[1] &&
                                         (unsigned)(u2 + t.lo[2]) < (unsigned)t.T[2];
                        const uint32_t ref = (uint32_t)(((b * t.L[0] + x0) * t.L[1] + x1) * t.L[2] + x2);
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
            for (int c = threadIdx.x; c < tcells; c += blockDim.x) {
                const int v0 = fdiv(c, trow, r_tr), rc = c - v0 * trow;
                const int v1 = fdiv(rc, t.T[2], r_T2), v2 = rc - v1 * t.T[2];
                const int u0 = v0 - t.lo[0], u1 = v1 - t.lo[1], u2 = v2 - t.lo[2];
                const int li = b * t.boxcells + (u0 * t.B[1] + u1) * t.B[2] + u2;
                // g0 + u is inside the lattice for the tile's own cells
                const int d = dev(b, g0[0] + u0, g0[1] + u1, g0[2] + u2);
                rec_st<R>(gout, (uint32_t)d * RB, X.ld(li));
            }
        }
        (void)r_T1;
        __syncthreads();  // the next tile's box overwrites LDS
    }
    if constexpr (KIND == 2) {
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0 && acc) atomicAdd(&a.acc[sys], (uint32_t)acc);
    }
}

template <class R, bool SELF, int KIND>
static hipError_t launch_tile_t(dim3 grid, size_t lds, const TileArgs &a, hipStream_t st) {
    if (lds > 65536) {
        const hipError_t e = allow_large_lds((const void *)k_tile<R, SELF, KIND>);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_tile<R, SELF, KIND>), grid, dim3(256), lds, st, a);
    return hipGetLastError();
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
