// ssf.hip -- equal-time spin correlations in momentum space
// (compute_equal_time_correlations, spin_correlations.jl:6-43) for every
// system of a batch at once.
//
// For momenta k and sites i at positions r_i,
//   s_a(k) = sum_i exp(-i k.r_i) s_{i,a},   S^{ab}(k) = Re(s_a(k) conj(s_b(k))) / N.
// With C = cos(k.r) and Q = sin(k.r) as (n_k x N) matrices and the spins of
// all R systems as one (N x 3R) matrix B, Re s = C B and Im s = -Q B, and
// S^{ab} = (Re_a Re_b + Im_a Im_b) / N (the sign of Im cancels).
//
// Default path, k_ssf_mfma: one FP64 MFMA kernel.  Each wave owns 16
// momenta and makes their phases in registers (FP64 sine and cosine of the
// FP64 argument, or rotations of the previous phase where the sites are
// regularly spaced; one per lane per 4 sites: the A operand of both
// products), the
// spins are staged through LDS in chunks of SSF_KC sites shared by the 4 waves
// (the B operand), and the N sum is split over workgroups to fill the
// device; a small kernel sums the partial products in a fixed order and forms
// the 9 correlations.  Nothing of size n_k x N ever reaches memory.
//
// The alternative path (the phase matrices written to memory and two
// rocBLAS dgemm per block of momenta) is kept for comparison.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "device.h"
#include "kernels.h"

#include <algorithm>

namespace csmc {

// C[kk + kb i] = cos(k.r_i), Q[kk + kb i] = sin(k.r_i) for momenta k0 .. k0+kb-1
// (column-major kb x N); pos and ks hold 3 components per entry
__global__ void k_ssf_phase(const double *pos, int N, const double *ks, int k0, int kb, double *C,
                            double *Q) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)kb * N) return;
    const int kk = (int)(t % kb), i = (int)(t / kb);
    const double *k = ks + 3 * (int64_t)(k0 + kk), *r = pos + 3 * (int64_t)i;
    const double arg = k[0] * r[0] + k[1] * r[1] + k[2] * r[2];
    double s, c;
    sincos(arg, &s, &c);
    C[t] = c;
    Q[t] = s;
}

// B[i + N (3 r + a)] = component a of site i (device order) of system r
template <class R>
__global__ void k_ssf_spins(const void *spins, const SysDev *sys, int nsys, int N, double *B) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nsys * N) return;
    const int r = (int)(t / N), i = (int)(t % N);
    const spin4<R> s = spins_at<R>(spins, sys[r].spin_off).ld(i);
    B[i + (int64_t)N * (3 * r + 0)] = (double)s.x;
    B[i + (int64_t)N * (3 * r + 1)] = (double)s.y;
    B[i + (int64_t)N * (3 * r + 2)] = (double)s.z;
}

// out[r][k][3 a + b] = (Re_a Re_b + Im_a Im_b) / N for k in the block (the
// reference's 9 x n_k column-major Suv per system)
__global__ void k_ssf_combine(const double *Cre, const double *Cim, int kb, int k0, int nk, int nsys,
                              double inv_n, double *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)kb * nsys) return;
    const int kk = (int)(t % kb), r = (int)(t / kb);
    double re[3], im[3];
    for (int a = 0; a < 3; ++a) {
        re[a] = Cre[kk + (int64_t)kb * (3 * r + a)];
        im[a] = Cim[kk + (int64_t)kb * (3 * r + a)];
    }
    double *o = out + ((int64_t)r * nk + k0 + kk) * 9;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) o[3 * a + b] = (re[a] * re[b] + im[a] * im[b]) * inv_n;
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// B[i * ldb + c] = component c % 3 of device site perm[i] of system c / 3;
// zero for i >= N or c >= 3 nsys (the padding the MFMA tiles read)
template <class R>
__global__ void k_ssf_spins_rows(const void *spins, const SysDev *sys, const int32_t *perm, int nsys, int N,
                                 int npad, int ldb, double *B) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)npad * ldb) return;
    const int i = (int)(t / ldb), c = (int)(t % ldb);
    double v = 0.0;
    if (i < N && c < 3 * nsys) {
        const spin4<R> s = spins_at<R>(spins, sys[c / 3].spin_off).ld(perm[i]);
        const int a = c % 3;
        v = (double)(a == 0 ? s.x : a == 1 ? s.y : s.z);
    }
    B[t] = v;
}

// sin and cos of x for |x| <= SSF_FAST_ARG (the host checks the bound):
// x = n pi/2 + r with pi/2 split in three parts (Cody and Waite; the first two
// have 33 significant bits, so n P1 and n P2 are exact for |n| < 2^20), the
// Taylor series of sin to r^15 and of cos to r^16 on |r| <= pi/4 (truncation
// below 5e-17), and the quadrant n mod 4.  No branch, so the phase of the
// next sites can be scheduled between the MFMAs of the current ones.
__device__ __forceinline__ void sincos_reduced(double x, double *sn, double *cs) {
    const double n = __builtin_rint(x * 0x1.45f306dc9c883p-1);  // 2 / pi
    double r = __builtin_fma(-n, 0x1.921fb54400000p+0, x);
    r = __builtin_fma(-n, 0x1.0b4611a600000p-34, r);
    r = __builtin_fma(-n, 0x1.3198a2e037073p-69, r);
    const double z = r * r;
    double sp = -1.0 / 1307674368000.0;  // -1 / 15!
    sp = __builtin_fma(sp, z, 1.0 / 6227020800.0);
    sp = __builtin_fma(sp, z, -1.0 / 39916800.0);
    sp = __builtin_fma(sp, z, 1.0 / 362880.0);
    sp = __builtin_fma(sp, z, -1.0 / 5040.0);
    sp = __builtin_fma(sp, z, 1.0 / 120.0);
    sp = __builtin_fma(sp, z, -1.0 / 6.0);
    const double s = __builtin_fma(r * z, sp, r);
    double cp = 1.0 / 20922789888000.0;  // 1 / 16!
    cp = __builtin_fma(cp, z, -1.0 / 87178291200.0);
    cp = __builtin_fma(cp, z, 1.0 / 479001600.0);
    cp = __builtin_fma(cp, z, -1.0 / 3628800.0);
    cp = __builtin_fma(cp, z, 1.0 / 40320.0);
    cp = __builtin_fma(cp, z, -1.0 / 720.0);
    cp = __builtin_fma(cp, z, 1.0 / 24.0);
    cp = __builtin_fma(cp, z, -0.5);
    const double c = __builtin_fma(z, cp, 1.0);
    const int q = (int)n & 3;
    const double sv = (q & 1) ? c : s, cv = (q & 1) ? s : c;
    *sn = (q & 2) ? -sv : sv;
    *cs = ((q + 1) & 2) ? -cv : cv;
}

// The branch-free sincos in six stages (the same operations in the same
// order as sincos_reduced), so a caller can place them between MFMAs:
// stage 0 takes x, stages 1-4 carry n, r, z and the two polynomials,
// phase_finish applies the quadrant.
__device__ __forceinline__ void phase_stage(int seg, double x, double &n, double &r, double &z, double &sp,
                                            double &cp, double &sv, double &cv) {
    if (seg == 0) {
        n = __builtin_rint(x * 0x1.45f306dc9c883p-1);
        r = __builtin_fma(-n, 0x1.921fb54400000p+0, x);
    } else if (seg == 1) {
        r = __builtin_fma(-n, 0x1.0b4611a600000p-34, r);
        r = __builtin_fma(-n, 0x1.3198a2e037073p-69, r);
        z = r * r;
    } else if (seg == 2) {
        sp = __builtin_fma(-1.0 / 1307674368000.0, z, 1.0 / 6227020800.0);
        sp = __builtin_fma(sp, z, -1.0 / 39916800.0);
        sp = __builtin_fma(sp, z, 1.0 / 362880.0);
        cp = __builtin_fma(1.0 / 20922789888000.0, z, -1.0 / 87178291200.0);
        cp = __builtin_fma(cp, z, 1.0 / 479001600.0);
        cp = __builtin_fma(cp, z, -1.0 / 3628800.0);
    } else if (seg == 3) {
        sp = __builtin_fma(sp, z, -1.0 / 5040.0);
        sp = __builtin_fma(sp, z, 1.0 / 120.0);
        sp = __builtin_fma(sp, z, -1.0 / 6.0);
        cp = __builtin_fma(cp, z, 1.0 / 40320.0);
        cp = __builtin_fma(cp, z, -1.0 / 720.0);
    } else if (seg == 4) {
        cp = __builtin_fma(cp, z, 1.0 / 24.0);
        cp = __builtin_fma(cp, z, -0.5);
        sv = __builtin_fma(r * z, sp, r);
        cv = __builtin_fma(z, cp, 1.0);
    }
}

__device__ __forceinline__ void phase_finish(double n, double sv, double cv, double *sn, double *cs) {
    const int q = (int)n & 3;
    const double s1 = (q & 1) ? cv : sv, c1 = (q & 1) ? sv : cv;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

// Partial products of a block of 64 momenta (16 per wave) x 16 NCT columns
// over the chunks [ch0, ch1) of SSF_KC sites:
//   P[((z 2 + 0) mpad + m) ldb + c] = sum_i cos(k_m.r_i) B[i][c]
//   P[((z 2 + 1) mpad + m) ldb + c] = sum_i sin(k_m.r_i) B[i][c]
// v_mfma_f64_16x16x4f64 operands: lane l gives A[l % 16][l / 16] (momentum
// l % 16 of the wave's 16, site l / 16 of the 4) and B[l / 16][l % 16];
// it holds D[l / 16 + 4 v][l % 16] in element v of its accumulator.
// pos (npad x 3), ks (mpad x 3) and B (npad x ldb) are zero-padded, so no
// load needs a bound check; padded sites carry zero spins.
//
// The phases of the lane's 4 sites at a time (PH):
//   0  the library sincos for every 4 sites;
//   1  the branch-free sincos for every 4 sites, its stages placed between
//      the MFMAs of the previous 4 (all |k.r| < SSF_FAST_ARG);
//   2  as 1, and every site 4 places later in the chunk is 4 places further
//      along the same vector d (the host checked): the phases of sites 4,
//      8 and 12 of a chunk are the previous ones rotated by exp(i k.d) (two
//      multiplies and two multiply-adds), and the branch-free sincos of the
//      next chunk's first 4 sites runs between the last MFMAs of the chunk.
//
// Work is spread as units (tile, chunk), a tile being 64 momenta x 16 NCT
// columns: workgroup g of the grid takes the units [g U / G, (g + 1) U / G)
// in tile-major order, so every workgroup has the same number of chunks and
// the grid is one round of resident workgroups.  Where its range crosses a
// tile boundary a workgroup writes one partial sum per tile it touched, to
// part g - wfirst[tile] of that tile (P: [tile][part][2][64][16 NCT]).
template <int NCT, int PH>
__global__ __launch_bounds__(256) void k_ssf_mfma(const double *__restrict__ pos, const double *__restrict__ ks,
                                                  const double *__restrict__ B, int ldb, int nchunks,
                                                  int colblocks, int maxparts, int64_t units,
                                                  const int32_t *__restrict__ wfirst, double dx, double dy,
                                                  double dz, double *__restrict__ P) {
    constexpr int KC = SSF_KC, BN = 16 * NCT, NK = KC / 4, NST = KC * BN / 256;
    static_assert(KC * BN % 256 == 0, "a chunk is a whole number of elements per thread");
    __shared__ double Bs[2][KC][BN];
    __shared__ double Ps[2][KC][3];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int64_t u_end = (int64_t)(blockIdx.x + 1) * units / gridDim.x;
    for (int64_t u = (int64_t)blockIdx.x * units / gridDim.x; u < u_end;) {
    const int tile = (int)(u / nchunks);
    const int ch0 = (int)(u % nchunks);
    const int ch1 = (int)min<int64_t>(nchunks, ch0 + (u_end - u));
    const int m = (tile / colblocks) * 64 + w * 16 + (l & 15);
    const double kx = ks[3 * m], ky = ks[3 * m + 1], kz = ks[3 * m + 2];
    const int c0 = (tile % colblocks) * BN;
    dbl4 ac[NCT], as[NCT];
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
        ac[t] = dbl4{0.0, 0.0, 0.0, 0.0};
        as[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    }
    // staging: element e = tid + 256 q of a chunk is row e / BN, column e % BN
    double stage[NST];
    double pstage = 0.0;
    auto load = [&](int ch) __attribute__((always_inline)) {
        const double *src = B + (int64_t)ch * KC * ldb + c0;
#pragma unroll
        for (int q = 0; q < NST; ++q) {
            const int e = tid + 256 * q;
            stage[q] = src[(int64_t)(e / BN) * ldb + e % BN];
        }
        if (PH != 2 && tid < 3 * KC) pstage = pos[(int64_t)ch * KC * 3 + tid];
    };
    auto store = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < NST; ++q) {
            const int e = tid + 256 * q;
            Bs[buf][e / BN][e % BN] = stage[q];
        }
        if (PH != 2 && tid < 3 * KC) Ps[buf][tid / 3][tid % 3] = pstage;
    };
    // the argument k.r of the lane's site among the first 4 of chunk ch
    auto first_arg = [&](int ch) __attribute__((always_inline)) {
        const double *pp = pos + ((int64_t)ch * KC + (l >> 4)) * 3;
        return kx * pp[0] + ky * pp[1] + kz * pp[2];
    };
    double cw = 1.0, sw = 0.0;  // exp(i k.d)
    double cs = 0.0, sn = 0.0;  // PH 2: the phase of the lane's current site
    if constexpr (PH == 2) {
        sincos_reduced(kx * dx + ky * dy + kz * dz, &sw, &cw);
        sincos_reduced(first_arg(ch0), &sn, &cs);
    }
    load(ch0);
    store(0);
    __syncthreads();
    for (int ch = ch0; ch < ch1; ++ch) {
        const int buf = (ch - ch0) & 1;
        if (ch + 1 < ch1) load(ch + 1);
        double bf[NCT];
        if constexpr (PH == 2) {
#pragma unroll
            for (int t = 0; t < NCT; ++t) bf[t] = Bs[buf][l >> 4][16 * t + (l & 15)];
            // the last chunk computes a phase it does not use (of a chunk in
            // bounds)
            const double xn = first_arg(min(ch + 1, nchunks - 1));
            double cs_n = 0.0, sn_n = 0.0;
#pragma unroll
            for (int kk = 0; kk < NK; ++kk) {
                const bool next = kk + 1 < NK;
                const int row = 4 * (kk + 1) + (l >> 4);
                double bn[NCT], n = 0.0, r = 0.0, z = 0.0, sp = 0.0, cp = 0.0, sv = 0.0, cv = 0.0;
#pragma unroll
                for (int seg = 0; seg < 6; ++seg) {
#pragma unroll
                    for (int t = seg * NCT / 6; t < (seg + 1) * NCT / 6; ++t) {
                        ac[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(cs, bf[t], ac[t], 0, 0, 0);
                        as[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, bf[t], as[t], 0, 0, 0);
                        if (next) bn[t] = Bs[buf][row][16 * t + (l & 15)];
                    }
                    if (!next) phase_stage(seg, xn, n, r, z, sp, cp, sv, cv);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (next) {
                    const double c1 = __builtin_fma(cs, cw, -(sn * sw));
                    const double s1 = __builtin_fma(sn, cw, cs * sw);
                    cs = c1;
                    sn = s1;
#pragma unroll
                    for (int t = 0; t < NCT; ++t) bf[t] = bn[t];
                } else {
                    phase_finish(n, sv, cv, &sn_n, &cs_n);
                }
            }
            cs = cs_n;
            sn = sn_n;
        } else if constexpr (PH == 1) {
            {
                const int row = l >> 4;
#pragma unroll
                for (int t = 0; t < NCT; ++t) bf[t] = Bs[buf][row][16 * t + (l & 15)];
                sincos_reduced(kx * Ps[buf][row][0] + ky * Ps[buf][row][1] + kz * Ps[buf][row][2], &sn, &cs);
            }
#pragma unroll
            for (int kk = 0; kk < NK; ++kk) {
                // six segments, each with the MFMAs of NCT / 6 column tiles
                // of these 4 sites, one stage of the phase of the next 4 and
                // the reads of their B fragments; the scheduling barriers
                // keep the stages between the MFMAs
                const bool next = kk + 1 < NK;
                const int row = 4 * (kk + 1) + (l >> 4);
                double bn[NCT], x = 0.0, n = 0.0, r = 0.0, z = 0.0, sp = 0.0, cp = 0.0, sv = 0.0, cv = 0.0;
#pragma unroll
                for (int seg = 0; seg < 6; ++seg) {
#pragma unroll
                    for (int t = seg * NCT / 6; t < (seg + 1) * NCT / 6; ++t) {
                        ac[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(cs, bf[t], ac[t], 0, 0, 0);
                        as[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, bf[t], as[t], 0, 0, 0);
                        if (next) bn[t] = Bs[buf][row][16 * t + (l & 15)];
                    }
                    if (next) {
                        if (seg == 0) x = kx * Ps[buf][row][0] + ky * Ps[buf][row][1] + kz * Ps[buf][row][2];
                        phase_stage(seg, x, n, r, z, sp, cp, sv, cv);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (next) {
                    phase_finish(n, sv, cv, &sn, &cs);
#pragma unroll
                    for (int t = 0; t < NCT; ++t) bf[t] = bn[t];
                }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < NK; ++kk) {
                const int row = 4 * kk + (l >> 4);
#pragma unroll
                for (int t = 0; t < NCT; ++t) bf[t] = Bs[buf][row][16 * t + (l & 15)];
                sincos(kx * Ps[buf][row][0] + ky * Ps[buf][row][1] + kz * Ps[buf][row][2], &sn, &cs);
#pragma unroll
                for (int t = 0; t < NCT; ++t) {
                    ac[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(cs, bf[t], ac[t], 0, 0, 0);
                    as[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sn, bf[t], as[t], 0, 0, 0);
                }
            }
        }
        if (ch + 1 < ch1) store(buf ^ 1);
        __syncthreads();
    }
    double *pc = P + ((int64_t)tile * maxparts + ((int)blockIdx.x - wfirst[tile])) * (2 * 64 * BN) +
                 (w * 16 + (l >> 4)) * BN + (l & 15);
    double *ps = pc + 64 * BN;
#pragma unroll
    for (int t = 0; t < NCT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            pc[4 * v * BN + 16 * t] = ac[t][v];
            ps[4 * v * BN + 16 * t] = as[t][v];
        }
    u += ch1 - ch0;
    __syncthreads();  // the next tile's staging reuses the LDS buffers
    }
}

// out[r][k][3 a + b] = (Re_a Re_b + Im_a Im_b) / N, the partial sums of the
// tile added in part order; consecutive threads take consecutive systems, so
// the partial sums are read in order of memory.  nparts: parts per tile.
__global__ void k_ssf_finish(const double *P, int colblocks, int bn, int maxparts, const int32_t *nparts, int nk,
                             int nsys, double inv_n, double *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nk * nsys) return;
    const int r = (int)(t % nsys), k = (int)(t / nsys);
    const int cb = 3 * r / bn, col = 3 * r - cb * bn;  // a system's 3 columns share a column block
    const int tile = (k / 64) * colblocks + cb;
    double re[3] = {0.0, 0.0, 0.0}, im[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < nparts[tile]; ++p) {
        const double *pc = P + ((int64_t)tile * maxparts + p) * (2 * 64 * bn) + (k % 64) * bn + col;
        const double *ps = pc + 64 * bn;
        for (int a = 0; a < 3; ++a) {
            re[a] += pc[a];
            im[a] += ps[a];
        }
    }
    double *o = out + ((int64_t)r * nk + k) * 9;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) o[3 * a + b] = (re[a] * re[b] + im[a] * im[b]) * inv_n;
}

template <int NCT>
static const void *mfma_kernel(int ph) {
    return ph == 2   ? (const void *)k_ssf_mfma<NCT, 2>
           : ph == 1 ? (const void *)k_ssf_mfma<NCT, 1>
                     : (const void *)k_ssf_mfma<NCT, 0>;
}

static const void *mfma_kernel(int nct, int ph) {
    switch (nct) {
    case 1: return mfma_kernel<1>(ph);
    case 2: return mfma_kernel<2>(ph);
    case 3: return mfma_kernel<3>(ph);
    case 4: return mfma_kernel<4>(ph);
    case 6: return mfma_kernel<6>(ph);
    case 8: return mfma_kernel<8>(ph);
    default: return mfma_kernel<12>(ph);
    }
}

void ssf_plan(int N, int nk, int nsys, int ncu, int ph, SsfPlan *p) {
    static const int ncts[] = {1, 2, 3, 4, 6, 8, 12};
    const int cols = 3 * nsys;
    int nct = 12;
    for (int c : ncts)
        if (16 * c >= cols) {
            nct = c;
            break;
        }
    p->N = N;
    p->nk = nk;
    p->nsys = nsys;
    p->nct = nct;
    p->ph = ph;
    p->colblocks = (cols + 16 * nct - 1) / (16 * nct);
    p->ldb = p->colblocks * 16 * nct;
    p->mblocks = (nk + 63) / 64;
    p->mpad = p->mblocks * 64;
    p->nchunks = (N + SSF_KC - 1) / SSF_KC;
    p->npad = p->nchunks * SSF_KC;
    // one round of resident workgroups, each with an equal range of units
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mfma_kernel(nct, ph), 256, 0) != hipSuccess || occ < 1)
        occ = 1;
    const int64_t tiles = (int64_t)p->mblocks * p->colblocks;
    p->units = tiles * p->nchunks;
    p->grid = (int)std::min<int64_t>(p->units, (int64_t)(ncu > 0 ? ncu : 256) * occ);
    // the workgroup holding unit u is the last g with g U / G <= u
    auto owner = [&](int64_t u) { return (int)(((u + 1) * p->grid - 1) / p->units); };
    p->wfirst.assign(tiles, 0);
    p->nparts.assign(tiles, 0);
    p->maxparts = 1;
    for (int64_t t = 0; t < tiles; ++t) {
        p->wfirst[t] = owner(t * p->nchunks);
        p->nparts[t] = owner((t + 1) * p->nchunks - 1) - p->wfirst[t] + 1;
        p->maxparts = std::max(p->maxparts, p->nparts[t]);
    }
}

hipError_t launch_ssf_mfma(const SsfPlan &p, int prec, const void *spins, const SysDev *sys, const int32_t *perm,
                           const int32_t *tiles, const double *pos, const double *ks, double *B, double *P,
                           double *out, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    const int64_t nb = (int64_t)p.npad * p.ldb;
    if (prec == 32)
        hipLaunchKernelGGL(k_ssf_spins_rows<float>, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, spins,
                           sys, perm, p.nsys, p.N, p.npad, p.ldb, B);
    else
        hipLaunchKernelGGL(k_ssf_spins_rows<double>, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st,
                           spins, sys, perm, p.nsys, p.N, p.npad, p.ldb, B);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    {
        int ldb = p.ldb, nchunks = p.nchunks, colblocks = p.colblocks, maxparts = p.maxparts;
        int64_t units = p.units;
        double dx = p.d[0], dy = p.d[1], dz = p.d[2];
        void *args[] = {(void *)&pos, (void *)&ks, (void *)&B, &ldb, &nchunks, &colblocks, &maxparts, &units,
                        (void *)&tiles, &dx, &dy, &dz, (void *)&P};
        e = hipLaunchKernel(mfma_kernel(p.nct, p.ph), dim3(p.grid), dim3(256), args, 0, st);
        if (e != hipSuccess) return e;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    const int64_t nc = (int64_t)p.nk * p.nsys;
    const int32_t *nparts = tiles + p.mblocks * p.colblocks;
    hipLaunchKernelGGL(k_ssf_finish, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, P, p.colblocks,
                       16 * p.nct, p.maxparts, nparts, p.nk, p.nsys, 1.0 / p.N, out);
    return hipGetLastError();
}

__global__ void k_ssf_accumulate(const double *out, const int32_t *sys_of_slot, int R, int nk, double *acc) {
    const int64_t per = (int64_t)nk * 9;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * R) return;
    const int T = (int)(t / per);
    acc[t] += out[(int64_t)sys_of_slot[T] * per + t % per];
}

hipError_t launch_ssf_accumulate(const double *out, const int32_t *sys_of_slot, int R, int nk, double *acc,
                                 hipStream_t st) {
    const int64_t n = (int64_t)R * nk * 9;
    hipLaunchKernelGGL(k_ssf_accumulate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, sys_of_slot, R,
                       nk, acc);
    return hipGetLastError();
}

static hipError_t blas_err(rocblas_status s) {
    return s == rocblas_status_success ? hipSuccess : hipErrorLaunchFailure;
}

hipError_t launch_ssf_blas(void *blas, int prec, const void *spins, const SysDev *sys, int nsys, int N,
                      const double *pos, const double *ks, int nk, int kb, double *C, double *Q,
                      double *B, double *Cre, double *Cim, double *out, hipStream_t st) {
    rocblas_handle h = (rocblas_handle)blas;
    hipError_t e = blas_err(rocblas_set_stream(h, st));
    if (e != hipSuccess) return e;
    const int64_t nb = (int64_t)nsys * N;
    if (prec == 32)
        hipLaunchKernelGGL(k_ssf_spins<float>, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, spins,
                           sys, nsys, N, B);
    else
        hipLaunchKernelGGL(k_ssf_spins<double>, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st,
                           spins, sys, nsys, N, B);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const double one = 1.0, zero = 0.0;
    for (int k0 = 0; k0 < nk; k0 += kb) {
        const int m = nk - k0 < kb ? nk - k0 : kb;
        const int64_t np = (int64_t)m * N;
        hipLaunchKernelGGL(k_ssf_phase, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, pos, N, ks,
                           k0, m, C, Q);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // (m x N) x (N x 3R), column-major throughout
        e = blas_err(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, 3 * nsys, N, &one,
                                   C, m, B, N, &zero, Cre, m));
        if (e != hipSuccess) return e;
        e = blas_err(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, 3 * nsys, N, &one,
                                   Q, m, B, N, &zero, Cim, m));
        if (e != hipSuccess) return e;
        const int64_t nc = (int64_t)m * nsys;
        hipLaunchKernelGGL(k_ssf_combine, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, Cre, Cim, m,
                           k0, nk, nsys, 1.0 / N, out);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t blas_create(void **blas) {
    rocblas_handle h = nullptr;
    hipError_t e = blas_err(rocblas_create_handle(&h));
    *blas = h;
    return e;
}

void blas_destroy(void *blas) {
    if (blas) (void)rocblas_destroy_handle((rocblas_handle)blas);
}

}  // namespace csmc
