// device.h -- gfx950 device building blocks of libcsmc: spin storage views,
// Philox4x32-10, the local field / site energy of ClassicalSpinMC.jl's
// hamiltonian.jl and the single-site updates of its sweeps.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "csmc_internal.h"

namespace csmc {

template <class R>
struct alignas(4 * sizeof(R)) spin4 {
    R x, y, z, w;
};

// View of one lattice's device tables (see csmc_internal.h for the layout).
// Read-only table data through the constant address space: with a
// wave-uniform address the compiler emits scalar loads (and hoists them out
// of site loops), which a plain global pointer does not allow because the
// kernel also writes global memory.
template <class T>
__device__ __forceinline__ T ldc(const T *p) {
    return *(const __attribute__((address_space(4))) T *)(p);
}

// f(b) for a site of basis b.  When every active lane of the wave has the
// same basis, b is passed as a wave-uniform value, so the coupling loads
// indexed by it are scalar loads; otherwise each lane passes its own.
template <class F>
__device__ __forceinline__ void with_wave_basis(int b, F &&f) {
    const int b0 = __builtin_amdgcn_readfirstlane(b);
    if (__all(b == b0))
        f(b0);
    else
        f(b);
}

template <class R>
struct LatView {
    int N, nb, K2, K2P, K3, K4, ncol, has_self, flags;
    const int32_t *nbr2, *nbr3, *nbr4, *ref, *sub, *colrun, *pat3, *pat4;
    const int4 *runs;
    const R *onsite, *field, *c2, *c3, *c4;
    R S;
};

template <class R>
__device__ __forceinline__ LatView<R> lat_view(const LatDev &d, const int32_t *i32,
                                               const int4 *runs, const R *coef) {
    LatView<R> v;
    v.N = d.N;
    v.nb = d.nb;
    v.K2 = d.K2;
    v.K2P = d.K2P;
    v.K3 = d.K3;
    v.K4 = d.K4;
    v.ncol = d.ncol;
    v.has_self = d.has_self;
    v.flags = d.flags;
    v.nbr2 = i32 + d.nbr_off;
    v.nbr3 = v.nbr2 + (int64_t)d.K2P * d.N;
    v.nbr4 = v.nbr3 + (int64_t)d.K3 * 2 * d.N;
    v.ref = i32 + d.ref_off;
    v.sub = i32 + d.sub_off;
    v.colrun = i32 + d.colrun_off;
    v.pat3 = i32 + d.pat_off;
    v.pat4 = v.pat3 + d.nb * d.K3;
    v.runs = runs;
    const R *c = coef + d.coef_off;
    v.onsite = c;
    v.field = c + d.nb * 9;
    v.c2 = v.field + d.nb * 3;
    v.c3 = v.c2 + (int64_t)d.nb * d.K2P * 9;
    v.c4 = v.c3 + (int64_t)d.nb * d.K3 * 27;
    v.S = (R)d.S;
    return v;
}

// ---------------------------------------------------------------------------
// spin storage views: packed three-component records in global memory (12 B
// per FP32 site, 24 B per FP64 site, no padding word), or three component
// arrays in LDS.  In registers a spin is a spin4 with w = 0.
// ---------------------------------------------------------------------------
template <class R>
struct rec3 {
    R x, y, z;
};

template <class R>
struct SpinsG {
    rec3<R> *p;
    __device__ __forceinline__ spin4<R> ld(int64_t i) const {
        const rec3<R> r = p[i];
        return spin4<R>{r.x, r.y, r.z, R(0)};
    }
    __device__ __forceinline__ void st(int64_t i, const spin4<R> &v) const { p[i] = rec3<R>{v.x, v.y, v.z}; }
};

// the records of the system whose first record is `off` in a spin buffer
template <class R>
__device__ __forceinline__ SpinsG<R> spins_at(void *buf, int64_t off) {
    return SpinsG<R>{reinterpret_cast<rec3<R> *>(buf) + off};
}
template <class R>
__device__ __forceinline__ SpinsG<R> spins_at(const void *buf, int64_t off) {
    return SpinsG<R>{const_cast<rec3<R> *>(reinterpret_cast<const rec3<R> *>(buf)) + off};
}

template <class R>
struct SpinsSoA {
    R *x, *y, *z;
    __device__ __forceinline__ spin4<R> ld(int i) const {
        spin4<R> s;
        s.x = x[i];
        s.y = y[i];
        s.z = z[i];
        s.w = R(0);
        return s;
    }
    __device__ __forceinline__ void st(int i, const spin4<R> &v) const {
        x[i] = v.x;
        y[i] = v.y;
        z[i] = v.z;
    }
};

// bilinear neighbour tables, slot-major, an absent neighbour is the zero pad
// site N: int32 in global memory with all K2P slots, or uint16 in LDS with
// the K2 real slots only (the padding slots then resolve to N without a load)
// at(k, i) is the load for a slot known to be real (k < K2).
struct NbrG {
    const int32_t *p;
    int N;
    __device__ __forceinline__ int operator()(int k, int i) const { return p[(int64_t)k * N + i]; }
    __device__ __forceinline__ int at(int k, int i) const { return p[(int64_t)k * N + i]; }
};

// The partners of the sites one thread visits in a sweep, staged in LDS as
// uint16 (lattices of up to 65535 sites) when the sites a thread updates stay
// the same for a whole launch: row s (bilinear slots 0..K2-1, then the two
// partners of each cubic slot and the three of each quartic slot) holds at
// column `slot` the partner of the site visited at that slot.  Serves the
// cubic and quartic partners too (kGen).
struct NbrPart {
    static constexpr bool kGen = true;
    const uint16_t *tab;
    int stride, slot, K2, K3, N;
    __device__ __forceinline__ int operator()(int k, int i) const { return k < K2 ? at(k, i) : N; }
    __device__ __forceinline__ int at(int k, int) const { return tab[k * stride + slot]; }
    __device__ __forceinline__ int c3(int k, int q, int) const { return tab[(K2 + 2 * k + q) * stride + slot]; }
    __device__ __forceinline__ int q4(int k, int q, int) const {
        return tab[(K2 + 2 * K3 + 3 * k + q) * stride + slot];
    }
};

// whether a neighbour functor serves the cubic and quartic partners (else
// they come from the lattice's global tables)
template <class T, class = void>
struct nb_has_gen : std::false_type {};
template <class T>
struct nb_has_gen<T, std::void_t<decltype(T::kGen)>> : std::bool_constant<T::kGen> {};

// whether a neighbour functor knows the lattice's bilinear slot count (its
// member K2): the field then skips the padding slots past it, which hold
// zero couplings and name the zero site (each would add +0 to the field)
template <class T, class = void>
struct nb_knows_k2 : std::false_type {};
template <class T>
struct nb_knows_k2<T, std::void_t<decltype(T::kKnowsK2)>> : std::bool_constant<T::kKnowsK2> {};

struct NbrL {
    const uint16_t *p;
    int N, K2;
    __device__ __forceinline__ int operator()(int k, int i) const {
        return k < K2 ? (int)p[k * N + i] : N;
    }
    __device__ __forceinline__ int at(int k, int i) const { return (int)p[k * N + i]; }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32x32->64 multiply per product (v_mad_u64_u32) gives both halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
                       (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
    }
    return c;
}

// uniforms in (0,1)
__device__ __forceinline__ double u01d(uint32_t x) {
    return ((double)x + 0.5) * 2.3283064365386963e-10;
}
__device__ __forceinline__ float u01f(uint32_t x) {
    return ((float)(x >> 8) + 0.5f) * 5.9604644775390625e-08f;
}

template <class R>
struct Math;

template <>
struct Math<double> {
    static __device__ __forceinline__ double u01(uint32_t x) { return u01d(x); }
    // random_spin_orientation (lattice.jl:306-311)
    static __device__ __forceinline__ void rand_spin(double S, double u1, double u2, double *o) {
        double phi = 2.0 * M_PI * u1;
        double z = 2.0 * u2 - 1.0;
        double r = sqrt(1.0 - z * z);
        double sn, cs;
        sincos(phi, &sn, &cs);
        o[0] = S * (r * cs);
        o[1] = S * (r * sn);
        o[2] = S * z;
    }
    // Boltzmann factor exp(-dE/T), evaluated as the reference writes it
    static __device__ __forceinline__ double boltz(double dE, double T, double) {
        return exp(-dE / T);
    }
    static __device__ __forceinline__ double rsqrt_(double x) { return 1.0 / sqrt(x); }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
};

template <>
struct Math<float> {
    static __device__ __forceinline__ float u01(uint32_t x) { return u01f(x); }
    // same formula; v_sin/v_cos take the angle in turns, so phi = 2 pi u1
    // needs no multiply and no range reduction; v_sqrt_f32 (1 ulp) serves
    // the FP32 state
    static __device__ __forceinline__ void rand_spin(float S, float u1, float u2, float *o) {
        float z = 2.0f * u2 - 1.0f;
        float r = __builtin_amdgcn_sqrtf(fmaxf(1.0f - z * z, 0.0f));
        o[0] = S * (r * __builtin_amdgcn_cosf(u1));
        o[1] = S * (r * __builtin_amdgcn_sinf(u1));
        o[2] = S * z;
    }
    static __device__ __forceinline__ float boltz(float dE, float, float invT) {
        return __expf(-dE * invT);
    }
    static __device__ __forceinline__ float rsqrt_(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float div(float a, float b) {
        float r = __builtin_amdgcn_rcpf(b);
        r = r * (2.0f - b * r);  // one Newton step
        return a * r;
    }
};

// ---------------------------------------------------------------------------
// Interaction field F_i = sum_n J_n s_j + sum C s_j s_k + sum R s_j s_k s_l
// (the non-onsite, non-Zeeman part of get_local_field, hamiltonian.jl:24-65).
// `nb(k, i)` gives the bilinear neighbours, `ld(j)` the spin of device site j;
// basis b is the site's basis (wave-uniform in the sweep kernels).  Bilinear
// slots are taken four at a time with every index load issued before the
// first use, so one memory round trip serves four neighbours; absent and
// padding slots hit the zero pad site, so there is no branch.  Couplings are
// read through the constant address space (scalar loads when b is uniform).
// ---------------------------------------------------------------------------
// Bilinear part of the field from the K2C neighbour spins of a site of basis
// b, as one multiply-add chain per component.  DIAG: every coupling matrix of
// the lattice is diagonal, so only m11, m22, m33 take part.
template <class R, class A, int K2C, bool DIAG, class S>
__device__ __forceinline__ void bilinear_field(const LatView<R> &L, int b, const S *sj, A F[3]) {
#pragma unroll
    for (int k = 0; k < K2C; ++k) {
        const R *M = L.c2 + (b * L.K2P + k) * 9;
        const A x = sj[k].x, y = sj[k].y, z = sj[k].z;
        if constexpr (DIAG) {
            if (k == 0) {
                F[0] = (A)ldc(M + 0) * x;
                F[1] = (A)ldc(M + 4) * y;
                F[2] = (A)ldc(M + 8) * z;
            } else {
                F[0] += (A)ldc(M + 0) * x;
                F[1] += (A)ldc(M + 4) * y;
                F[2] += (A)ldc(M + 8) * z;
            }
        } else {
            if (k == 0) {
                F[0] = (A)ldc(M + 0) * x;
                F[1] = (A)ldc(M + 3) * x;
                F[2] = (A)ldc(M + 6) * x;
            } else {
                F[0] += (A)ldc(M + 0) * x;
                F[1] += (A)ldc(M + 3) * x;
                F[2] += (A)ldc(M + 6) * x;
            }
            F[0] += (A)ldc(M + 1) * y;
            F[1] += (A)ldc(M + 4) * y;
            F[2] += (A)ldc(M + 7) * y;
            F[0] += (A)ldc(M + 2) * z;
            F[1] += (A)ldc(M + 5) * z;
            F[2] += (A)ldc(M + 8) * z;
        }
    }
}

// Entries a sparsity pattern keeps (TP_* of csmc_internal.h); p is the
// site's own component, q, r, t index the partners' spins.
__device__ constexpr bool c3_kept(int pat, int p, int q, int r) {
    return pat == TP_DENSE || (p != q && q != r && p != r);
}
__device__ constexpr bool q4_kept(int pat, int p, int q, int r, int t) {
    return pat == TP_DENSE || (pat == TP_Q4_PQ && p == q && r == t) ||
           (pat == TP_Q4_PR && p == r && q == t) || (pat == TP_Q4_PT && p == t && q == r);
}

// F_p += sum_q (sum_r C[p,q,r] sk_r) sj_q for one cubic slot
// (hamiltonian.jl:36-49, at the site's position).  Every loop is
// unrolled, so the pattern tests are resolved at compile time and the
// products kept are added in the order of the full contraction.
template <int PAT, class R, class A>
__device__ __forceinline__ void cubic_term(const R *C, const A sj[3], const A sk[3], A F[3]) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        A acc = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (!(c3_kept(PAT, p, q, 0) || c3_kept(PAT, p, q, 1) || c3_kept(PAT, p, q, 2))) continue;
            A t = 0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (c3_kept(PAT, p, q, r)) t += (A)ldc(C + p + 3 * q + 9 * r) * sk[r];
            acc += t * sj[q];
        }
        F[p] += acc;
    }
}

// F_p += sum_q (sum_{r,t} Q[p,q,r,t] sk_r sl_t) sj_q for one quartic slot
// (hamiltonian.jl:51-65), with the pattern's entries only
template <int PAT, class R, class A>
__device__ __forceinline__ void quartic_term(const R *Q, const A sj[3], const A sk[3], const A sl[3],
                                             A F[3]) {
    A w[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int t = 0; t < 3; ++t) w[r + 3 * t] = sk[r] * sl[t];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        A acc = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            bool any = false;
#pragma unroll
            for (int rt = 0; rt < 9; ++rt) any = any || q4_kept(PAT, p, q, rt % 3, rt / 3);
            if (!any) continue;
            A t = 0;
#pragma unroll
            for (int rt = 0; rt < 9; ++rt)
                if (q4_kept(PAT, p, q, rt % 3, rt / 3)) t += (A)ldc(Q + p + 3 * q + 9 * rt) * w[rt];
            acc += t * sj[q];
        }
        F[p] += acc;
    }
}

// The coefficient base address made opaque to the optimiser at each call:
// loads through it cannot be hoisted out of the caller's site loop.  With
// cubic and quartic slots, hoisting every slot's coefficients into scalar
// registers for a whole run of sites takes hundreds of SGPRs (spilled to
// vector registers and then to scratch); reloading a slot's few kept
// coefficients per site with scalar loads is cheaper.  Emits no instruction.
template <class T>
__device__ __forceinline__ const T *per_site_coef(const T *p) {
    const uint64_t v = (uint64_t)(size_t)p;
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    asm volatile("" : "+s"(lo), "+s"(hi));
    return (const T *)(size_t)(((uint64_t)hi << 32) | lo);
}

// K2C > 0: the lattice has exactly K2C bilinear slots (compile-time), so
// the slot loop is fully unrolled with no padding work.
// PART 0: the whole field; 1: its bilinear and cubic part; 2: add the
// quartic part to F (0 is 1 followed by 2, with the same operations).
// SP: every cubic and quartic slot of the lattice has a sparse pattern (no
// TP_DENSE slot), so the full contractions are not compiled in: they would
// hold 27 or 81 coefficients per slot in registers.
template <class R, class A, bool GEN = true, int K2C = 0, int PART = 0, bool SP = false, class NB,
          class LD>
__device__ __forceinline__ void interaction_field(const LatView<R> &L, const NB &nb, int b, int i,
                                                  LD &&ld, A F[3]) {
    if (PART != 2) F[0] = F[1] = F[2] = A(0);
    // lattices with cubic or quartic terms reload coefficients per site
    const R *const c2 = GEN ? per_site_coef(L.c2) : L.c2;
    if constexpr (K2C > 0 && PART != 2) {
        int j[K2C];
#pragma unroll
        for (int k = 0; k < K2C; ++k) j[k] = nb.at(k, i);
        spin4<R> sj[K2C];
#pragma unroll
        for (int k = 0; k < K2C; ++k) sj[k] = ld(j[k]);
        bilinear_field<R, A, K2C, false>(L, b, sj, F);
    }
    // the bilinear slots in groups of four; DIAG: every coupling matrix is
    // diagonal, and each row below rounds to m_rr s_r (its other terms are
    // +-0), so only those products are formed, each rounded before the sum
    // (a product that cannot be fused into the addition)
    auto slots = [&](auto diag) {
        constexpr bool DIAG = decltype(diag)::value;
        for (int k0 = 0; k0 < L.K2P; k0 += 4) {
            int ku = 4;
            if constexpr (nb_knows_k2<NB>::value) ku = min(4, nb.K2 - k0);
            int j[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) j[u] = nb(k0 + u, i);
            spin4<R> sj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) sj[u] = ld(j[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (u >= ku) break;
                const R *M = c2 + (b * L.K2P + k0 + u) * 9;
                if constexpr (DIAG) {
#pragma clang fp contract(off)
                    F[0] += (A)ldc(M + 0) * sj[u].x;
                    F[1] += (A)ldc(M + 4) * sj[u].y;
                    F[2] += (A)ldc(M + 8) * sj[u].z;
                } else {
                    // each row as fma(m2, z, fma(m0, x, m1 y)), written out
                    // so that every kernel rounds it alike
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                        F[r] += fma((A)ldc(M + 3 * r + 2), (A)sj[u].z,
                                    fma((A)ldc(M + 3 * r), (A)sj[u].x, (A)ldc(M + 3 * r + 1) * (A)sj[u].y));
                }
            }
        }
    };
    if constexpr (K2C == 0 && PART != 2) {
        if (nb_knows_k2<NB>::value && (L.flags & LAT_DIAG))  // wave-uniform
            slots(std::true_type{});
        else
            slots(std::false_type{});
    }
    if (!GEN) return;
    // GS slots at a time: their patterns, partner indices and partner spins
    // are loaded before any of them is used, so the loads of a group are in
    // flight together (a group past the last slot repeats the last slot's
    // loads and skips its product).  With the partner indices in LDS
    // (NbrPart) a group of three measured faster; with indices gathered from
    // global memory, two.
    constexpr int GS = nb_has_gen<NB>::value ? 3 : 2;
    for (int k0 = 0; PART != 2 && k0 < L.K3; k0 += GS) {
        int pat[GS], j[GS][2];
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            const int k = min(k0 + u, L.K3 - 1);
            const int p = ldc(L.pat3 + b * L.K3 + k);
            pat[u] = k0 + u < L.K3 ? p : TP_ZERO;
            if constexpr (nb_has_gen<NB>::value) {
                j[u][0] = nb.c3(k, 0, i);
                j[u][1] = nb.c3(k, 1, i);
            } else {
                j[u][0] = L.nbr3[((int64_t)k * 2) * L.N + i];
                j[u][1] = L.nbr3[((int64_t)k * 2 + 1) * L.N + i];
            }
        }
        spin4<R> sp[GS][2];
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            sp[u][0] = ld(j[u][0]);
            sp[u][1] = ld(j[u][1]);
        }
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            if (pat[u] == TP_ZERO) continue;
            const A sj[3] = {sp[u][0].x, sp[u][0].y, sp[u][0].z}, sk[3] = {sp[u][1].x, sp[u][1].y, sp[u][1].z};
            const R *C = per_site_coef(L.c3) + (b * L.K3 + k0 + u) * 27;
            if (SP || pat[u] == TP_C3_ALT)
                cubic_term<TP_C3_ALT, R>(C, sj, sk, F);
            else
                cubic_term<TP_DENSE, R>(C, sj, sk, F);
        }
    }
    if (PART == 1) return;
    for (int k0 = 0; k0 < L.K4; k0 += GS) {
        int pat[GS], j[GS][3];
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            const int k = min(k0 + u, L.K4 - 1);
            const int p = ldc(L.pat4 + b * L.K4 + k);
            pat[u] = k0 + u < L.K4 ? p : TP_ZERO;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if constexpr (nb_has_gen<NB>::value)
                    j[u][q] = nb.q4(k, q, i);
                else
                    j[u][q] = L.nbr4[((int64_t)k * 3 + q) * L.N + i];
            }
        }
        spin4<R> sp[GS][3];
#pragma unroll
        for (int u = 0; u < GS; ++u)
#pragma unroll
            for (int q = 0; q < 3; ++q) sp[u][q] = ld(j[u][q]);
#pragma unroll
        for (int u = 0; u < GS; ++u) {
            if (pat[u] == TP_ZERO) continue;
            const A sj[3] = {sp[u][0].x, sp[u][0].y, sp[u][0].z}, sk[3] = {sp[u][1].x, sp[u][1].y, sp[u][1].z},
                    sl[3] = {sp[u][2].x, sp[u][2].y, sp[u][2].z};
            const R *Q = per_site_coef(L.c4) + (b * L.K4 + k0 + u) * 81;
            if (pat[u] == TP_Q4_PQ)
                quartic_term<TP_Q4_PQ, R>(Q, sj, sk, sl, F);
            else if (pat[u] == TP_Q4_PR)
                quartic_term<TP_Q4_PR, R>(Q, sj, sk, sl, F);
            else if (SP || pat[u] == TP_Q4_PT)
                quartic_term<TP_Q4_PT, R>(Q, sj, sk, sl, F);
            else
                quartic_term<TP_DENSE, R>(Q, sj, sk, sl, F);
        }
    }
}

// get_local_field (hamiltonian.jl:3-67): H = 2 O s + F - h
// (the onsite and Zeeman parts are skipped by a uniform branch when the
// lattice has none: 2 O s and h are then exactly zero)
template <class R, class A>
__device__ __forceinline__ void local_field(const LatView<R> &L, int b, const spin4<R> &s,
                                            const A F[3], A H[3]) {
    const R *O = L.onsite + b * 9;
    const R *h = L.field + b * 3;
    A sx = s.x, sy = s.y, sz = s.z;
    H[0] = F[0];
    H[1] = F[1];
    H[2] = F[2];
    if (L.flags & LAT_ONSITE) {
        H[0] = A(2) * ((A)ldc(O + 0) * sx + (A)ldc(O + 1) * sy + (A)ldc(O + 2) * sz) + H[0];
        H[1] = A(2) * ((A)ldc(O + 3) * sx + (A)ldc(O + 4) * sy + (A)ldc(O + 5) * sz) + H[1];
        H[2] = A(2) * ((A)ldc(O + 6) * sx + (A)ldc(O + 7) * sy + (A)ldc(O + 8) * sz) + H[2];
    }
    if (L.flags & LAT_FIELD) {
        H[0] -= (A)ldc(h + 0);
        H[1] -= (A)ldc(h + 1);
        H[2] -= (A)ldc(h + 2);
    }
}

// s.O.s
template <class R, class A>
__device__ __forceinline__ A onsite_energy(const LatView<R> &L, int b, A sx, A sy, A sz) {
    if (!(L.flags & LAT_ONSITE)) return A(0);
    const R *O = L.onsite + b * 9;
    return sx * ((A)ldc(O + 0) * sx + (A)ldc(O + 1) * sy + (A)ldc(O + 2) * sz) +
           sy * ((A)ldc(O + 3) * sx + (A)ldc(O + 4) * sy + (A)ldc(O + 5) * sz) +
           sz * ((A)ldc(O + 6) * sx + (A)ldc(O + 7) * sy + (A)ldc(O + 8) * sz);
}

// ---------------------------------------------------------------------------
// single-site updates; X is a spin view, nb the bilinear neighbour view
// ---------------------------------------------------------------------------

// overrelaxation! body (monte_carlo.jl:128-137) for a site with spin s and
// local field H: false (no write) when H == 0
template <class R>
__device__ __forceinline__ bool or_reflect(const spin4<R> &s, const R H[3], spin4<R> &n) {
    if (H[0] == R(0) && H[1] == R(0) && H[2] == R(0)) return false;
    R hh = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
    R proj = Math<R>::div(R(2) * (s.x * H[0] + s.y * H[1] + s.z * H[2]), hh);
    n.x = -s.x + proj * H[0];
    n.y = -s.y + proj * H[1];
    n.z = -s.z + proj * H[2];
    n.w = R(0);
    return true;
}

// deterministic_updates! body (monte_carlo.jl:203-209): s = -S H / |H|
template <class R>
__device__ __forceinline__ bool det_align(const LatView<R> &L, const R H[3], spin4<R> &n) {
    if (H[0] == R(0) && H[1] == R(0) && H[2] == R(0)) return false;
    R nrm = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    n.x = -H[0] / nrm * L.S;
    n.y = -H[1] / nrm * L.S;
    n.z = -H[2] / nrm * L.S;
    n.w = R(0);
    return true;
}

template <class R, bool GEN, int K2C, bool SP = false, class NB, class SV>
__device__ __forceinline__ void or_update(const LatView<R> &L, const NB &nb, const SV &X, int b,
                                          int i) {
    spin4<R> s = X.ld(i), n;
    R F[3], H[3];
    interaction_field<R, R, GEN, K2C, 0, SP>(L, nb, b, i, [&](int j) { return X.ld(j); }, F);
    local_field<R, R>(L, b, s, F, H);
    if (or_reflect<R>(s, H, n)) X.st(i, n);
}

template <class R, bool GEN, int K2C, bool SP = false, class NB, class SV>
__device__ __forceinline__ void det_update(const LatView<R> &L, const NB &nb, const SV &X, int b,
                                           int i) {
    spin4<R> s = X.ld(i), n;
    R F[3], H[3];
    interaction_field<R, R, GEN, K2C, 0, SP>(L, nb, b, i, [&](int j) { return X.ld(j); }, F);
    local_field<R, R>(L, b, s, F, H);
    if (det_align<R>(L, H, n)) X.st(i, n);
}

// One Metropolis proposal (metropolis.jl:70-78 with calculate_energy_diff!,
// metropolis.jl:94-110).  kind: CSMC_METROPOLIS draws a fresh direction,
// the cone kinds draw s + sigma*r renormalised (gaussian_move,
// metropolis.jl:84-87).  Counter = (reference site, stream, ctr, METR).
// SELF: some term of this lattice names the site itself, so the new energy
// is evaluated with the candidate substituted in those slots as the
// reference does (it writes the spin before calling energy()).
// The pieces of one Metropolis step, shared by every sweep form.
// Proposal from the draws r.x, r.y: a fresh direction (CSMC_METROPOLIS) or
// s + sigma * (uniform direction) renormalised (gaussian_move,
// metropolis.jl:84-87).
template <class R>
__device__ __forceinline__ spin4<R> metro_propose(const LatView<R> &L, const spin4<R> &s, uint4 r,
                                                  int kind, R sigma) {
    R rs[3];
    Math<R>::rand_spin(L.S, Math<R>::u01(r.x), Math<R>::u01(r.y), rs);
    spin4<R> n;
    if (kind == 1) {
        n.x = rs[0];
        n.y = rs[1];
        n.z = rs[2];
    } else {
        R a = s.x + sigma * rs[0], c = s.y + sigma * rs[1], d = s.z + sigma * rs[2];
        R inv = Math<R>::rsqrt_(a * a + c * c + d * d) * L.S;
        n.x = a * inv;
        n.y = c * inv;
        n.z = d * inv;
    }
    n.w = R(0);
    return n;
}

// energy(new) - energy(old) at a site whose terms do not name the site
// itself: onsite difference plus (n - s).(F - h)
template <class R>
__device__ __forceinline__ R metro_dE(const LatView<R> &L, int b, const spin4<R> &s,
                                      const spin4<R> &n, const R F[3]) {
    const R *hp = L.field + b * 3;
    const R h0 = ldc(hp), h1 = ldc(hp + 1), h2 = ldc(hp + 2);
    R eo_new = onsite_energy<R, R>(L, b, n.x, n.y, n.z);
    R eo_old = onsite_energy<R, R>(L, b, s.x, s.y, s.z);
    return (eo_new - eo_old) + (n.x - s.x) * (F[0] - h0) + (n.y - s.y) * (F[1] - h1) +
           (n.z - s.z) * (F[2] - h2);
}

// accept if dE < 0, else if u < exp(-dE/T) (metropolis.jl:75-78), u = r.z
template <class R>
__device__ __forceinline__ int metro_accept(R dE, uint4 r, R T, R invT) {
    return (dE < R(0)) ? 1 : (Math<R>::u01(r.z) < Math<R>::boltz(dE, T, invT));
}

// store the accepted proposal; an FP32 state also rewrites a rejected spin
// pulled back to |s| = S, so reflections cannot accumulate rounding drift
template <class R, class SV, class IX>
__device__ __forceinline__ void metro_commit(const LatView<R> &L, const SV &X, IX i, int acc,
                                             spin4<R> s, const spin4<R> &n) {
    if (acc) {
        X.st(i, n);
    } else if (sizeof(R) == 4) {
        R inv = Math<R>::rsqrt_(s.x * s.x + s.y * s.y + s.z * s.z) * L.S;
        s.x *= inv;
        s.y *= inv;
        s.z *= inv;
        X.st(i, s);
    }
}

// energy(new) - energy(old) of the proposal n for the spin s at site i
// (calculate_energy_diff!, metropolis.jl:94-110)
template <class R, bool SELF, bool GEN, int K2C, bool SP = false, class NB, class SV>
__device__ __forceinline__ R metro_eval(const LatView<R> &L, const NB &nb, const SV &X, int b, int i,
                                        const spin4<R> &s, const spin4<R> &n) {
    R dE;
    if (!SELF) {
        R F[3];
        interaction_field<R, R, GEN, K2C, 0, SP>(L, nb, b, i, [&](int j) { return X.ld(j); }, F);
        dE = metro_dE<R>(L, b, s, n, F);
    } else {
        const R *hp = L.field + b * 3;
        const R h[3] = {ldc(hp), ldc(hp + 1), ldc(hp + 2)};
        R Fo[3], Fn[3];
        auto ldo = [&](int j) { return X.ld(j); };
        auto ldn = [&](int j) { return j == i ? n : X.ld(j); };
        if (GEN && (L.flags & LAT_SELF_Q4)) {
            // only quartic slots name the site: the bilinear and cubic part
            // is the same for both spins, the quartic part is added twice
            interaction_field<R, R, GEN, K2C, 1, SP>(L, nb, b, i, ldo, Fo);
            Fn[0] = Fo[0];
            Fn[1] = Fo[1];
            Fn[2] = Fo[2];
            interaction_field<R, R, GEN, K2C, 2, SP>(L, nb, b, i, ldo, Fo);
            interaction_field<R, R, GEN, K2C, 2, SP>(L, nb, b, i, ldn, Fn);
        } else {
            interaction_field<R, R, GEN, K2C, 0, SP>(L, nb, b, i, ldo, Fo);
            interaction_field<R, R, GEN, K2C, 0, SP>(L, nb, b, i, ldn, Fn);
        }
        R eo = onsite_energy<R, R>(L, b, s.x, s.y, s.z) + s.x * Fo[0] + s.y * Fo[1] + s.z * Fo[2] -
               (s.x * h[0] + s.y * h[1] + s.z * h[2]);
        R en = onsite_energy<R, R>(L, b, n.x, n.y, n.z) + n.x * Fn[0] + n.y * Fn[1] + n.z * Fn[2] -
               (n.x * h[0] + n.y * h[1] + n.z * h[2]);
        dE = en - eo;
    }
    return dE;
}

template <class R, bool SELF, bool GEN, int K2C, bool SP = false, class NB, class SV>
__device__ __forceinline__ int metro_update(const LatView<R> &L, const NB &nb, const SV &X, int b,
                                            int i, R T, R invT, int kind, R sigma, uint32_t k0,
                                            uint32_t k1, uint32_t stream, uint32_t ctr) {
    spin4<R> s = X.ld(i);
    uint4 r = philox4x32_10(make_uint4((uint32_t)L.ref[i], stream, ctr, TAG_METRO), k0, k1);
    spin4<R> n = metro_propose<R>(L, s, r, kind, sigma);
    const R dE = metro_eval<R, SELF, GEN, K2C, SP>(L, nb, X, b, i, s, n);
    int acc = metro_accept<R>(dE, r, T, invT);
    metro_commit<R>(L, X, i, acc, s, n);
    return acc;
}

// ---------------------------------------------------------------------------
// wave / workgroup reductions (wave64 shuffles)
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// DPP row shift of a double (both halves): lane i of each row of 16 reads
// lane i - k of its row, and 0 where that lane lies outside the row.
template <int K>
__device__ __forceinline__ double dpp_row_shr(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x110 + K, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + K, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// inclusive prefix sum within each row of 16 lanes: lane 15 of a row holds
// the row's sum, added in a fixed order
__device__ __forceinline__ double row_sum16(double v) {
    v += dpp_row_shr<1>(v);
    v += dpp_row_shr<2>(v);
    v += dpp_row_shr<4>(v);
    v += dpp_row_shr<8>(v);
    return v;
}

__device__ __forceinline__ double read_lane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sums NV <= 9 values per thread over the workgroup (every thread active, at
// most 16 waves); the totals are valid in every thread on return.  scratch:
// 153 doubles of LDS (per-wave sums in [0, 16 NV), totals in [144, 144 + NV)),
// not touched by other code between calls.
//
// Within a wave: DPP prefix sums over each row of 16 lanes, then the four
// row sums (read with readlane) added in row order -- VALU latency only, no
// trip through the LDS permute unit.  Across waves: wave 0 loads the per-wave
// sums, 16 waves of one value to a row, and sums each row the same way.
// Every operation and its order is fixed, so the result is the same on every
// run.  The totals lie past every call's per-wave sums and are rewritten
// only after the next call's first barrier, so no barrier is needed after
// they are read.
constexpr int BLOCK_SUM_SCRATCH = 16 * 9 + 9;

template <int NV>
__device__ __forceinline__ void block_sum(double *v, double *scratch) {
    static_assert(NV >= 1 && NV <= 9, "block_sum: at most 9 values");
    constexpr int TOT = 16 * 9;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double r = row_sum16(v[k]);
        const double t = ((read_lane_d(r, 15) + read_lane_d(r, 31)) + read_lane_d(r, 47)) +
                         read_lane_d(r, 63);
        if (lane == 0) scratch[wave * NV + k] = t;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int r0 = 0; r0 < NV; r0 += 4) {
            const int k = r0 + (lane >> 4), w = lane & 15;
            double x = (k < NV && w < nw) ? scratch[w * NV + k] : 0.0;
            x = row_sum16(x);
            if (w == 15 && k < NV) scratch[TOT + k] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = scratch[TOT + k];
}

}  // namespace csmc
