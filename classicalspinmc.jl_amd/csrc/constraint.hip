// constraint.hip -- the Lagrange-penalty Metropolis sweep of
// metropolis_constraint_adaptive! (metropolis.jl:155-179) for a constraint
// the device can evaluate: a sum over sites of a per-basis linear and
// quadratic term,
//     c(lattice) = sum_i (a_b . s_i + s_i . Q_b s_i) - c0,   b = basis of i.
//
// Per proposal the reference draws a cone move at a site, evaluates the
// constraint on the lattice holding the new spin, sets
// new_lambda = lambda - weight * c and accepts on dE - new_lambda * c
// (metropolis.jl:160-175); lambda changes with every accepted proposal, so
// the decisions form one chain per system.  Within a colour class the sites
// share no term, so each proposal's energy change and constraint change
//     delta_c = f_b(new) - f_b(old)
// do not depend on the other proposals of the class.  A colour pass is
// therefore three launches:
//   k_cons_propose  every site of the class in parallel: the proposal, dE,
//                   delta_c and the acceptance uniform;
//   k_cons_scan     one thread per system walks the class in site order:
//                   c += delta_c, new_lambda, the FP64 decision, and on
//                   acceptance c and lambda move on -- exactly the reference's
//                   sequence over the class;
//   k_cons_commit   every site of the class in parallel: accepted proposals
//                   are stored.
// k_cons_begin sums c over the system at the start of each sweep (the
// overrelaxation sweeps between Metropolis sweeps change it).  After a
// system's last colour the scan applies the width law of
// metropolis_constraint_adaptive! (metropolis.jl:176-177) and advances the
// system's Philox sweep counter.
#include "sweep.h"

namespace csmc {

// f_b(s) = a_b . s + s . Q_b s in FP64
__device__ __forceinline__ double cons_term(const ConsArgs &a, int b, double x, double y, double z) {
    if (b >= a.cnb) return 0.0;
    const double *q = a.coef + b * 12;
    const double lin = q[0] * x + q[1] * y + q[2] * z;
    const double qx = q[3] * x + q[4] * y + q[5] * z;
    const double qy = q[6] * x + q[7] * y + q[8] * z;
    const double qz = q[9] * x + q[10] * y + q[11] * z;
    return lin + (x * qx + y * qy + z * qz);
}

// c of every system at the start of a sweep; clears the sweep's count
template <class R>
__global__ __launch_bounds__(1024) void k_cons_begin(ConsArgs a) {
    __shared__ double red[BLOCK_SUM_SCRATCH];
    const int sys = blockIdx.x;
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    const int32_t *sub = a.P.i32 + ld.sub_off;
    const SpinsG<R> S = spins_at<R>(a.spins, sd.spin_off);
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < ld.N; i += blockDim.x) {
        const spin4<R> s = S.ld(i);
        v[0] += cons_term(a, sub[i], (double)s.x, (double)s.y, (double)s.z);
    }
    block_sum<1>(v, red);
    if (threadIdx.x == 0) {
        a.cval[sys] = v[0] - a.c0;
        a.acc_sweep[sys] = 0.0;
    }
}

template <class R, bool SELF, bool GEN>
__global__ __launch_bounds__(256) void k_cons_propose(ConsArgs a) {
    const int sys = blockIdx.y;
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    if (a.colour >= ld.ncol) return;
    const int4 *runs = a.P.runs + ld.run_off;
    const LatView<R> L = lat_view<R>(ld, a.P.i32, runs, coef_of<R>(a.P));
    const int start = runs[L.colrun[a.colour]].x;
    const int end = runs[L.colrun[a.colour + 1] - 1].y;
    const int i = start + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const SpinsG<R> X = spins_at<R>(a.spins, sd.spin_off);
    const NbrG nb{L.nbr2, L.N};
    const int b = L.sub[i];
    const spin4<R> s = X.ld(i);
    const uint4 r = philox4x32_10(make_uint4((uint32_t)L.ref[i], (uint32_t)sd.stream, a.ctr[sys], TAG_METRO),
                                  a.k0, a.k1);
    // the cone move of calculate_energy_diff!(lattice, point, sigma)
    // (metropolis.jl:102-110, gaussian_move :84-87)
    const spin4<R> n = metro_propose<R>(L, s, r, CSMC_CONE, (R)a.sigma_sys[sys]);
    const R dE = metro_eval<R, SELF, GEN, 0>(L, nb, X, b, i, s, n);
    const int64_t k = sd.spin_off + i;
    spins_at<R>(a.prop, 0).st(k, n);
    a.dE[k] = (double)dE;
    a.dC[k] = cons_term(a, b, (double)n.x, (double)n.y, (double)n.z) -
              cons_term(a, b, (double)s.x, (double)s.y, (double)s.z);
    a.u[k] = u01d(r.z);
}

// one thread per system: the reference's per-proposal chain over the class
__global__ void k_cons_scan(ConsArgs a, int nsys) {
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= nsys) return;
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    if (a.colour >= ld.ncol) return;
    const int4 *runs = a.P.runs + ld.run_off;
    const int32_t *colrun = a.P.i32 + ld.colrun_off;
    const int start = runs[colrun[a.colour]].x;
    const int end = runs[colrun[a.colour + 1] - 1].y;
    const double T = a.T_sys[sys], w = a.weight;
    double c = a.cval[sys], lam = a.lambda[sys], acc = 0.0;
    for (int i = start; i < end; ++i) {
        const int64_t k = sd.spin_off + i;
        const double cn = c + a.dC[k];
        const double ln = lam - w * cn;  // new_lambda = mc.lambda - mc.weight * c
        const double x = a.dE[k] - ln * cn;
        const int ok = (x < 0.0) ? 1 : (a.u[k] < exp(-x / T));
        a.dec[k] = (uint8_t)ok;
        if (ok) {
            c = cn;
            lam = ln;
            acc += 1.0;
        }
    }
    a.cval[sys] = c;
    a.lambda[sys] = lam;
    const double total = a.acc_sweep[sys] + acc;
    if (a.colour == ld.ncol - 1) {
        // metropolis.jl:176-177: f = 0.5 / (1 - rate); widen (or narrow) the
        // cone below 50 % acceptance, else reset it to sigma0
        const double rate = total / (double)ld.N;
        const double f = 0.5 / (1.0 - rate);
        a.sigma_sys[sys] = rate < 0.5 ? a.sigma_sys[sys] * f : a.sigma0;
        a.acc_total[sys] += total;
        a.acc_sweep[sys] = 0.0;
        a.ctr[sys] += 1u;
    } else {
        a.acc_sweep[sys] = total;
    }
}

template <class R>
__global__ __launch_bounds__(256) void k_cons_commit(ConsArgs a) {
    const int sys = blockIdx.y;
    const SysDev sd = a.P.sys[sys];
    const LatDev ld = a.P.lats[sd.lat];
    if (a.colour >= ld.ncol) return;
    const int4 *runs = a.P.runs + ld.run_off;
    const int32_t *colrun = a.P.i32 + ld.colrun_off;
    const int start = runs[colrun[a.colour]].x;
    const int end = runs[colrun[a.colour + 1] - 1].y;
    const int i = start + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const SpinsG<R> X = spins_at<R>(a.spins, sd.spin_off);
    const int64_t k = sd.spin_off + i;
    if (a.dec[k]) {
        X.st(i, spins_at<R>(a.prop, 0).ld(k));
    } else if (sizeof(R) == 4) {
        // an FP32 state pulls a rejected spin back to |s| = S (as the other
        // Metropolis kinds do)
        spin4<R> s = X.ld(i);
        const R inv = Math<R>::rsqrt_(s.x * s.x + s.y * s.y + s.z * s.z) * (R)ld.S;
        s.x *= inv;
        s.y *= inv;
        s.z *= inv;
        X.st(i, s);
    }
}

template <class R>
static void cons_pass_r(bool self, bool gen, int nsys, int max_colour_sites, const ConsArgs &a,
                        hipStream_t st) {
    dim3 grid((max_colour_sites + 255) / 256, nsys);
    if (gen) {
        if (self) hipLaunchKernelGGL((k_cons_propose<R, true, true>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_cons_propose<R, false, true>), grid, dim3(256), 0, st, a);
    } else {
        if (self) hipLaunchKernelGGL((k_cons_propose<R, true, false>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_cons_propose<R, false, false>), grid, dim3(256), 0, st, a);
    }
    hipLaunchKernelGGL(k_cons_scan, dim3((nsys + 63) / 64), dim3(64), 0, st, a, nsys);
    hipLaunchKernelGGL((k_cons_commit<R>), grid, dim3(256), 0, st, a);
}

hipError_t launch_cons_begin(int prec, int nsys, const ConsArgs &a, hipStream_t st) {
    if (prec == 32) hipLaunchKernelGGL((k_cons_begin<float>), dim3(nsys), dim3(1024), 0, st, a);
    else hipLaunchKernelGGL((k_cons_begin<double>), dim3(nsys), dim3(1024), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_cons_pass(int prec, bool self, bool gen, int nsys, int max_colour_sites, const ConsArgs &a,
                            hipStream_t st) {
    if (prec == 32) cons_pass_r<float>(self, gen, nsys, max_colour_sites, a, st);
    else cons_pass_r<double>(self, gen, nsys, max_colour_sites, a, st);
    return hipGetLastError();
}

}  // namespace csmc
