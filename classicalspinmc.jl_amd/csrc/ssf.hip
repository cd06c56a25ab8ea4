// ssf.hip -- equal-time spin correlations in momentum space
// (compute_equal_time_correlations, spin_correlations.jl:6-43) for every
// system of a batch at once.
//
// For momenta k and sites i at positions r_i,
//   s_a(k) = sum_i exp(-i k.r_i) s_{i,a},   S^{ab}(k) = Re(s_a(k) conj(s_b(k))) / N.
// With C = cos(k.r) and Q = sin(k.r) as (n_k x N) matrices and the spins of
// all R systems as one (N x 3R) matrix B, Re s = C B and Im s = -Q B: two
// FP64 GEMMs (rocBLAS; MFMA on gfx950) per block of momenta, so the phases
// are generated once for all systems.  S^{ab} = (Re_a Re_b + Im_a Im_b) / N
// (the sign of Im cancels).  Phases are FP64 sincos of the FP64 argument.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "device.h"
#include "kernels.h"

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
    const spin4<R> s = reinterpret_cast<const spin4<R> *>(spins)[sys[r].spin_off + i];
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

static hipError_t blas_err(rocblas_status s) {
    return s == rocblas_status_success ? hipSuccess : hipErrorLaunchFailure;
}

hipError_t launch_ssf(void *blas, int prec, const void *spins, const SysDev *sys, int nsys, int N,
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
