"""Thermal coupling derivatives d<E>/dJ = <dE/dJ> - (<E dE/dJ> - <E><dE/dJ>)/T
from Metropolis chains and the adjoint kernel, against the closed form of a
free spin in a field: <E> = -h S L(h S / T), L(x) = coth x - 1/x."""
import math

import numpy as np
import pytest

import csmc
from csmc.thermal import thermal_coupling_gradient

pytestmark = pytest.mark.gpu


def test_free_spin_thermal_field_derivative():
    h, S, T = 0.7, 1.0, 0.5
    uc = csmc.Square()
    csmc.addZeemanCoupling(uc, 1, [0.0, 0.0, h])
    lat = csmc.Lattice((1, 1), uc, S, seed=0)
    b = csmc.SpinBatch([lat] * 2048, precision=64)
    b.randomize(7)
    b.set_counters(0)
    g, err, means = thermal_coupling_gradient(b, T, n_samples=40, sweeps_between=4, thermalize=20, seed=3)
    x = h * S / T
    L = 1.0 / math.tanh(x) - 1.0 / x
    dL = 1.0 / x ** 2 - 1.0 / math.sinh(x) ** 2
    exact_z = -S * L - h * S * dL * S / T          # d<E>/dh_z
    assert abs(means["energy"] + h * S * L) < 5 * S * h * 0.02, means["energy"]
    assert abs(g[2] - exact_z) < 4 * err[2], (g[2], exact_z, err[2])
    # no field along x, y: the derivative is -<s_x> plus a covariance, both 0
    assert abs(g[0]) < 4 * err[0] and abs(g[1]) < 4 * err[1], (g[:2], err[:2])
    assert err[2] < 0.05 * abs(exact_z)


def _j2_direction(uc, n_j1=6, n_j2=6):
    """the direction of J2 in coupling space: J2 * I3 in the six J2 bilinear
    terms of triangular_j1j2 (9 entries per bilinear term, in call order)"""
    w = np.zeros(uc.n_params())
    for t in range(n_j1, n_j1 + n_j2):
        w[9 * t + np.array([0, 4, 8])] = 1.0
    return w


def test_triangular_j1j2_thermal_dE_dJ2_full_size():
    """The differentiation workload at full size: triangular J1-J2, L=96
    (9216 sites), S=1, J1=1, J2=0.1, 128 FP64 chains, T=0.1, 80 samples per
    chain (10,240 samples).  d<E>/dJ2 from the covariance estimator must
    agree within 4 sigma with a central finite difference of <E> itself,
    taken from two more runs at J2 +- 0.005 with the same seeds (paired per
    chain).  At T -> 0 the derivative tends to dE0/dJ2 of the 120-degree
    state, where every J2 pair is parallel: 3 N S^2 = 27,648 with this
    energy convention (each pair entered in both directions, E halved);
    thermal fluctuations raise it by a few per cent at T=0.1."""
    from csmc import workloads as W
    L, n, T, J2, d = 96, 128, 0.1, 0.1, 0.005

    def run(j2):
        uc = W.triangular_j1j2(1.0, j2)
        lat = csmc.Lattice((L, L), uc, 1.0, seed=5)
        b = csmc.SpinBatch([lat] * n, precision=64)
        b.randomize(7)
        b.set_counters(0)
        b.anneal(W.anneal_temperatures(1.0, 0.9, T), 100, 5, csmc.CONE_ADAPTIVE, seed=11)
        out = thermal_coupling_gradient(b, T, 80, sweeps_between=10, overrelaxation=5, thermalize=200,
                                        alg=csmc.CONE, sigma=0.3, seed=11, directions=_j2_direction(uc))
        G0 = None
        if j2 == J2:
            b.set_spins(0, np.tile([0.0, 0.0, 1.0], (L * L, 1)))
            G0 = float(b.grad_couplings()[0] @ _j2_direction(uc))
        b.close()
        return out, G0

    (g, err, means), G0 = run(J2)
    (_, _, mp), _ = run(J2 + d)
    (_, _, mm), _ = run(J2 - d)
    assert means["samples"] == n * 80
    fd = (mp["energy_per_system"] - mm["energy_per_system"]) / (2 * d)
    fd_mean, fd_err = fd.mean(), fd.std(ddof=1) / math.sqrt(n)
    assert G0 == 3 * L * L
    assert abs(g - fd_mean) < 4 * math.hypot(err, fd_err), (g, err, fd_mean, fd_err)
    assert err < 0.01 * G0 and fd_err < 0.01 * G0
    assert G0 < g < 1.05 * G0 and G0 < fd_mean < 1.05 * G0, (g, fd_mean, G0)
