// kernels.h -- launch interface between the libcsmc host runtime and its
// gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csmc_internal.h"

namespace csmc {

struct DevPools {
    const LatDev *lats;
    const int32_t *i32;
    const int4 *runs;
    const float *coefF;
    const double *coefD;
    const SysDev *sys;
};

enum WGMode { WG_SWEEPS = 0, WG_ANNEAL = 1, WG_PT = 2 };

// Arguments of the persistent workgroup-per-system kernel.
struct WGArgs {
    DevPools P;
    void *spins;
    int mode;
    int kind;           // sweep kind (WG_SWEEPS) or Metropolis alg (anneal / PT)
    uint32_t k0, k1;    // Philox key (seed)
    uint32_t *mcount;   // [nsys] Metropolis sweep counters
    // WG_SWEEPS
    int n_sweeps;
    const double *T_sys;
    double *sigma_sys;
    double *acc_sys;
    // WG_ANNEAL
    const double *temps;
    int n_temps;
    int64_t t_th, or_rate;
    double sigma0;
    double *acc_T;      // [nsys][n_temps]
    // WG_PT
    PTDev pt;
    int64_t seg_begin, seg_end;
    int begin_mid, end_mid, exchange;
};

// Arguments of the one-colour-pass kernel (plain / stream form).
struct PassArgs {
    DevPools P;
    void *spins;
    int sys0;
    int colour;
    int kind;
    uint32_t k0, k1;
    const uint32_t *ctr;     // [nsys] Philox sweep counter word per system, or null
    uint32_t ctr0;           // counter word used when ctr is null
    const double *T_sys;
    const double *sigma_sys;
    uint32_t *acc;           // [nsys] accepted proposals
};

int max_lds_bytes();
// lm: storage mode (0 global, 1 LDS spins, 2 LDS spins + uint16 neighbour
// tables); self / gen: some lattice of the batch has a self-referencing
// term / cubic or quartic terms; k2c: the bilinear slot count shared by
// every lattice of the batch (0 if they differ)
hipError_t launch_wg(int prec, int lm, bool self, bool gen, int k2c, int nsys, size_t lds_bytes,
                     const WGArgs &a, hipStream_t st);
hipError_t launch_pass(int prec, bool self, bool gen, int nsys, int max_colour_sites,
                       const PassArgs &a, hipStream_t st);
hipError_t launch_site_eval(int prec, const DevPools &P, const void *spins, int sys, int N,
                            double *H, double *e, hipStream_t st);
hipError_t launch_energy(int prec, const DevPools &P, const void *spins, int nsys, double *E,
                         double *parts, hipStream_t st);
hipError_t launch_magnet(int prec, const DevPools &P, const void *spins, int nsys, double *M,
                         hipStream_t st);
hipError_t launch_randomize(int prec, const DevPools &P, void *spins, int nsys, int maxN,
                            uint32_t k0, uint32_t k1, hipStream_t st);
// src: double[N][3] in reference order (device buffer)
hipError_t launch_scatter_spins(int prec, const DevPools &P, void *spins, int sys, int N,
                                const double *src, hipStream_t st);
hipError_t launch_gather_spins(int prec, const DevPools &P, const void *spins, int sys, int N,
                               double *dst, hipStream_t st);
// dEdS: double[total sites][3], systems concatenated in batch order
hipError_t launch_grad_spins(int prec, const DevPools &P, const void *spins, int nsys, int maxN,
                             double *dEdS, hipStream_t st);
hipError_t launch_grad_couplings(int prec, const DevPools &P, const void *spins, int nsys,
                                 int grad_len, double *dEdJ, hipStream_t st);
hipError_t launch_gd(int prec, const DevPools &P, void *spins, int nsys, int maxN,
                     const double *dEdS, double eta, hipStream_t st);

}  // namespace csmc
