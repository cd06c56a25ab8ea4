"""Independent lattices of uneven size in one launch (square and triangular
lattices with per-lattice Dzyaloshinskii-Moriya couplings and fields): every
system against the oracle, FP64 -- coloured overrelaxation to 1e-11,
Metropolis decisions identical with per-system temperatures and Philox
streams, and a two-temperature annealing schedule replayed step by step."""
import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from golden_cases import colour_order

pytestmark = pytest.mark.gpu


def _systems():
    # sizes chosen to give 2, 3 and 4 colours and K2 = 4 and 6 in one batch
    cells = W.uneven_batch(n=6, sizes=(5, 6, 8, 9, 12), seed=7)
    lats = [csmc.Lattice(shape, uc, 1.0, seed=20 + k) for k, (uc, shape) in enumerate(cells)]
    tabs = [orc.Tables(uc, shape) for uc, shape in cells]
    return lats, tabs


def _close(a, b, rtol):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) <= rtol * max(1.0, float(np.max(np.abs(b))))


def test_uneven_batch_sweeps_match_oracle():
    lats, tabs = _systems()
    assert len({lat.size for lat in lats}) > 2
    b = csmc.SpinBatch(lats, precision=64)
    spins = [lat._spins.copy() for lat in lats]
    for k, s in enumerate(spins):
        b.set_spins(k, s)
    orders = [colour_order(lat) for lat in lats]
    E = b.total_energy()
    for k in range(len(lats)):
        assert abs(E[k] - orc.total_energy(tabs[k], spins[k])) <= 1e-12 * lats[k].size
    b.sweep(csmc.OVERRELAX, n_sweeps=2)
    for k in range(len(lats)):
        for _ in range(2):
            orc.overrelax(tabs[k], spins[k], orders[k])
        assert _close(b.get_spins(k), spins[k], 1e-11), k
    temps = np.linspace(0.3, 1.5, len(lats))
    seed = 4242
    b.set_counters(3)
    acc = b.sweep(csmc.METROPOLIS, T=temps, n_sweeps=1, seed=seed)
    for k in range(len(lats)):
        a, _, _ = orc.metropolis_coloured(tabs[k], spins[k], orders[k], 1.0, temps[k], 0, 60.0, seed, k, 3)
        assert acc[k] == a, k
        assert _close(b.get_spins(k), spins[k], 1e-10), k
    b.close()


def test_uneven_batch_anneal_replay():
    # csmc_anneal: for each temperature, t = 1 .. t_th - 1 overrelaxation
    # sweeps with a Metropolis sweep when t % or_rate == 0
    # (monte_carlo.jl:166-187), every system in the same launch
    lats, tabs = _systems()
    b = csmc.SpinBatch(lats, precision=64)
    spins = [lat._spins.copy() for lat in lats]
    for k, s in enumerate(spins):
        b.set_spins(k, s)
    b.set_counters(0)
    orders = [colour_order(lat) for lat in lats]
    temps, t_th, or_rate, seed = [0.9, 0.5], 7, 3, 99
    acc = b.anneal(temps, t_th, or_rate, csmc.METROPOLIS, seed=seed)
    for k in range(len(lats)):
        ctr = 0
        for j, T in enumerate(temps):
            a_tot = 0
            for t in range(1, t_th):
                orc.overrelax(tabs[k], spins[k], orders[k])
                if t % or_rate == 0:
                    a, _, _ = orc.metropolis_coloured(tabs[k], spins[k], orders[k], 1.0, T, 0, 60.0, seed, k, ctr)
                    ctr += 1
                    a_tot += a
            assert acc[k, j] == a_tot, (k, j)
        assert _close(b.get_spins(k), spins[k], 1e-9), k
    b.close()


def test_uneven_batch_fp32_anneal_lowers_energy():
    # FP32 state, the configuration the batched anneal runs in: energies
    # fall with temperature and |s| stays S
    lats, _ = _systems()
    b = csmc.SpinBatch(lats, precision=32)
    b.randomize(5)
    E0 = b.total_energy() / np.array([lat.size for lat in lats])
    b.anneal(W.anneal_temperatures(1.0, 0.9, 0.05), 200, 10, csmc.METROPOLIS, seed=1)
    E1 = b.total_energy() / np.array([lat.size for lat in lats])
    assert np.all(E1 < E0 - 0.5)
    for k, lat in enumerate(lats):
        assert np.allclose(np.linalg.norm(b.get_spins(k), axis=1), 1.0, atol=1e-5)
    b.close()
