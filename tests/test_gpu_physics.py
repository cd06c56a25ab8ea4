"""Physics checks on the GPU through the reference-shaped API: the reference's
own known-answer test, closed forms, and statistical parity (3 sigma) with
the reference's random-site chain run by the oracle."""
import math

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from conftest import random_spins

pytestmark = pytest.mark.gpu


def _ground_state(uc, shape, S, t_th, Tmin, t_det, alg=None, precision=64, seed=11):
    lat = csmc.Lattice(shape, uc, S, seed=seed)
    params = {"t_thermalization": t_th, "overrelaxation_rate": 10, "t_deterministic": t_det}
    mc = csmc.MonteCarlo(Tmin, lat, params, precision=precision, seed=seed)
    csmc.simulated_annealing(mc, lambda x: 1.0 * 0.9 ** x, 1.0, alg=alg or csmc.Metropolis())
    if t_det > 1:
        csmc.deterministic_updates(mc)
    return mc


def test_kgg_ground_state_kat():
    # test/mctests.jl:43-50: annealing 0.9^x from 1 to 1e-7 (t_th = 1e4,
    # or_rate = 10) then deterministic updates; round(E/N, 4) == -0.6444
    mc = _ground_state(W.kgg_test_cell(), (4, 4), 1.0, 10 ** 4, 1e-7, 10 ** 6)
    assert round(csmc.energy_density(mc.lattice), 4) == -0.6444


def test_kgg_ground_state_kat_adaptive():
    # test/mctests.jl:52-58: MetropolisAdaptive, annealing only
    mc = _ground_state(W.kgg_test_cell(), (4, 4), 1.0, 10 ** 4, 1e-7, 1, alg=csmc.MetropolisAdaptive())
    assert round(csmc.energy_density(mc.lattice), 4) == -0.6444


def test_square_fm_ground_state():
    # FM along +z with h = 0.1 z: E/N = -2 S^2 |J| - h S = -2.1, |M| = N S
    mc = _ground_state(W.square_heisenberg(-1.0, (0, 0, 0.1)), (16, 16), 1.0, 2000, 1e-3, 2 * 10 ** 5)
    assert abs(csmc.energy_density(mc.lattice) + 2.1) < 1e-9
    assert abs(csmc.get_magnetization(mc.lattice) - 256) < 1e-6


@pytest.mark.parametrize("precision", [64, 32])
def test_klein_dual_ground_state(precision):
    # honeycomb J = 1, K = -2 is the Klein dual of the FM Heisenberg model
    # J' = -1: ground state E/N = -(3/2) S^2 = -0.375 at S = 1/2
    mc = _ground_state(W.honeycomb_kitaev(1.0, -2.0), (4, 4), 0.5, 2000, 1e-4, 10 ** 4,
                       precision=precision)
    tol = 1e-9 if precision == 64 else 1e-5
    assert abs(csmc.energy_density(mc.lattice) + 0.375) < tol


@pytest.mark.parametrize("precision", [64, 32])
def test_single_spin_langevin(precision):
    # a free spin in a field: <s.h^> = S L(beta h S), L(x) = coth x - 1/x
    uc = csmc.Square()
    csmc.addZeemanCoupling(uc, 1, [0.0, 0.0, 0.7])
    lat = csmc.Lattice((1, 1), uc, 1.0, seed=0)
    n = 4096
    b = csmc.SpinBatch([lat] * n, precision=precision)
    b.randomize(5)
    T = 0.5
    b.sweep(csmc.METROPOLIS, T=T, n_sweeps=20, seed=9)
    zs = []
    for k in range(40):
        b.sweep(csmc.METROPOLIS, T=T, n_sweeps=5, seed=9)
        zs.append(b.magnetization()[:, 2])
    zs = np.array(zs)  # (40, n): systems independent, samples 5 sweeps apart
    x = 0.7 / T
    exact = 1.0 / math.tanh(x) - 1.0 / x
    per_sys = zs.mean(axis=0)
    err = per_sys.std(ddof=1) / math.sqrt(n)
    assert abs(per_sys.mean() - exact) < 4 * err, (per_sys.mean(), exact, err)


def _oracle_chain(T, spins, temp, S, n_sweeps, or_rate, seed):
    """the reference's sweep schedule (monte_carlo.jl:295-305) with random-site
    Metropolis and sequential overrelaxation, returning E and |M| per
    Metropolis sweep"""
    order = np.arange(1, T.N + 1)
    Es, Ms = [], []
    for sw in range(n_sweeps):
        orc.overrelax(T, spins, order)
        if sw % or_rate == 0:
            orc.metropolis_random(T, spins, S, temp, 0, 60.0, seed, 0, sw)
            Es.append(orc.total_energy(T, spins))
            Ms.append(orc.magnetization(spins)[0])
    return np.array(Es), np.array(Ms)


@pytest.mark.parametrize("precision", [64, 32])
def test_statistical_parity_with_reference_chain(precision):
    # square Heisenberg FM L=6 at T = 1.2 (paramagnet, short correlation
    # time): <E> and <|M|> from the device PT observables vs the reference's
    # random-site chain, within 3 combined sigma; both sides use several
    # independent chains and take the error from the spread of chain means
    uc = W.square_heisenberg(-1.0, (0, 0, 0.1))
    L, S, temp = 6, 1.0, 1.2
    lat = csmc.Lattice((L, L), uc, S, seed=4)
    N = lat.size
    T = orc.Tables(uc, (L, L))
    n_sw, n_ch = 50000, 8
    chains_E, chains_M = [], []
    for c in range(n_ch):
        s = np.ascontiguousarray(lat._spins.copy())
        Eo, Mo = _oracle_chain(T, s, temp, S, n_sw, 2, 12345 + c)
        chains_E.append(Eo[1000:].mean())
        chains_M.append(Mo[1000:].mean())
    me_o, de_o = np.mean(chains_E), np.std(chains_E, ddof=1) / math.sqrt(n_ch)
    mm_o, dm_o = np.mean(chains_M), np.std(chains_M, ddof=1) / math.sqrt(n_ch)
    R = 8  # independent replicas at the same temperature (no exchange effect)
    mcs = [csmc.MonteCarlo(temp, lat, {"t_thermalization": 2000, "t_measurement": n_sw,
                                       "probe_rate": 2, "swap_rate": 2, "overrelaxation_rate": 2,
                                       "report_interval": 10 ** 9, "checkpoint_rate": 10 ** 9},
                           precision=precision, seed=77) for _ in range(R)]
    csmc.parallel_tempering(mcs, verbose=False)
    me = np.array([m.observables["energy"] for m in mcs])
    mm = np.array([m.observables["magnetization"] for m in mcs])
    me_g, de_g = me.mean(), me.std(ddof=1) / math.sqrt(R)
    mm_g, dm_g = mm.mean(), mm.std(ddof=1) / math.sqrt(R)
    assert abs(me_g - me_o) < 3 * math.hypot(de_g, de_o), (me_g / N, me_o / N, de_g, de_o)
    assert abs(mm_g - mm_o) < 3 * math.hypot(dm_g, dm_o), (mm_g / N, mm_o / N, dm_g, dm_o)


def test_pt_observables_match_host_binning():
    # the FP64 device binning of E and |M| (observables.jl:20-63) against
    # the oracle's binner fed with the same device trajectory
    uc = W.square_heisenberg(-1.0, (0, 0, 0.1))
    lat = csmc.Lattice((4, 4), uc, 1.0, seed=6)
    R = 4
    temps = W.geometric_temperatures(0.5, 2.0, R)
    p = dict(t_thermalization=6, t_measurement=600, probe_rate=3, swap_rate=2, overrelaxation_rate=2,
             report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    b = csmc.SpinBatch([lat] * R, precision=64)
    b.randomize(3)
    b.pt_init(temps, p, csmc.METROPOLIS, seed=5)
    binE = [orc.Binner() for _ in range(R)]
    binM = [orc.Binner() for _ in range(R)]
    total = p["t_thermalization"] + p["t_measurement"]
    for sw in range(total):
        b.pt_advance(sw + 1)
        if sw >= p["t_thermalization"] and sw % p["probe_rate"] == 0:
            st = b.pt_state()
            M = b.magnetization()
            for r in range(R):
                binE[r].push(st["E"][r])
                binM[r].push(float(np.linalg.norm(M[st["system_of_slot"][r]])))
    obs = b.pt_observables()
    for r in range(R):
        ref = orc.observables(binE[r], binM[r], temps[r], lat.size)
        got = [obs[k][r] for k in ("specific_heat", "specific_heat_err", "susceptibility",
                                   "susceptibility_err", "energy", "energy_err", "magnetization",
                                   "magnetization_err")]
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-12), (r, got, ref)
