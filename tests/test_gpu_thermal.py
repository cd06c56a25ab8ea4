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
