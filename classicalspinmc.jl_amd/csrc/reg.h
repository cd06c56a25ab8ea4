// reg.h -- pieces shared by the register-table sweep kernels (whole-system
// and split forms): float4 spin records in LDS addressed by absolute LDS byte
// address, and the bilinear couplings of one basis held in registers.
#pragma once
#include "sweep.h"

namespace csmc {

// spins as spin4<float> records in LDS, addressed by their absolute LDS byte
// address (so an access needs no base add)
typedef float lds_f4 __attribute__((ext_vector_type(4)));
using lds_byte = __attribute__((address_space(3))) unsigned char;
using lds_vec = __attribute__((address_space(3))) lds_f4;

__device__ __forceinline__ uint32_t lds_address(const void *p) {
    return (uint32_t)(size_t)(const lds_byte *)p;
}

// An LDS address held as one opaque VGPR value.  Without this the compiler
// keeps base-relative offsets in the address tables and adds the (uniform)
// base back before every LDS access, one VALU add per access; an empty asm
// with a "+v" operand hides the sum, so the tables hold final addresses.
__device__ __forceinline__ uint32_t pin_address(uint32_t a) {
    asm volatile("" : "+v"(a));
    return a;
}

struct SpinsV4 {
    __device__ __forceinline__ spin4<float> ld(uint32_t a) const {
        const lds_f4 v = *(const lds_vec *)(size_t)a;
        return spin4<float>{v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ void st(uint32_t a, const spin4<float> &v) const {
        *(lds_vec *)(size_t)a = lds_f4{v.x, v.y, v.z, v.w};
    }
};

// the bilinear couplings of one basis, K2C slots, held in registers (the
// values are wave-uniform, so they live in SGPRs); DIAG keeps m11, m22, m33
template <class S, int K2C, bool DIAG>
struct BilCoef {
    static constexpr int W = DIAG ? 3 : 9;
    S m[K2C][W];
    __device__ __forceinline__ void load(const LatView<S> &L, int b) {
#pragma unroll
        for (int k = 0; k < K2C; ++k)
#pragma unroll
            for (int e = 0; e < W; ++e) m[k][e] = ldc(L.c2 + (b * L.K2P + k) * 9 + (DIAG ? 4 * e : e));
    }
    // the bilinear field, in the operation order of bilinear_field
    template <class A>
    __device__ __forceinline__ void field(const spin4<float> *sj, A F[3]) const {
#pragma unroll
        for (int k = 0; k < K2C; ++k) {
            const A x = sj[k].x, y = sj[k].y, z = sj[k].z;
            if constexpr (DIAG) {
                if (k == 0) {
                    F[0] = (A)m[k][0] * x;
                    F[1] = (A)m[k][1] * y;
                    F[2] = (A)m[k][2] * z;
                } else {
                    F[0] += (A)m[k][0] * x;
                    F[1] += (A)m[k][1] * y;
                    F[2] += (A)m[k][2] * z;
                }
            } else {
                if (k == 0) {
                    F[0] = (A)m[k][0] * x;
                    F[1] = (A)m[k][3] * x;
                    F[2] = (A)m[k][6] * x;
                } else {
                    F[0] += (A)m[k][0] * x;
                    F[1] += (A)m[k][3] * x;
                    F[2] += (A)m[k][6] * x;
                }
                F[0] += (A)m[k][1] * y;
                F[1] += (A)m[k][4] * y;
                F[2] += (A)m[k][7] * y;
                F[0] += (A)m[k][2] * z;
                F[1] += (A)m[k][5] * z;
                F[2] += (A)m[k][8] * z;
            }
        }
    }
};

}  // namespace csmc
