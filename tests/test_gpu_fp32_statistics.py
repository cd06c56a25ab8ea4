"""Statistical parity of the FP32 PT path (the kernel the bench times) with
what the mathematics and the reference's own chain fix:

  - Klein duality: the honeycomb J=1, K=-2 model at S=1/2 is mapped onto the
    honeycomb ferromagnet J'=-1 by site-dependent pi rotations (a 2x2
    supercell pattern), which preserve the energy of every configuration.
    So the two models have the same energy distribution at every
    temperature: per temperature slot, <E> (observables.jl:20-25), the
    specific heat (observables.jl:36-49) and the replica-exchange acceptance
    (monte_carlo.jl:326-331, a function of the energy distributions of
    neighbouring slots) must agree.
  - The reference's sequence (monte_carlo.jl:295-349): random-site
    Metropolis, sequential overrelaxation, total_energy per Metropolis sweep
    and the even/odd exchange, run by the oracle: per slot <E>, C_v and the
    exchange acceptance of the device's coloured FP32 PT must agree with it.

Errors are the spread of independent runs (different seeds) on each side:
z = (mean_a - mean_b) / sqrt(sem_a^2 + sem_b^2), and every |z| must be
below 3.
"""
import math

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W

pytestmark = pytest.mark.gpu

Z_MAX = 3.0


def _z(a, b):
    a, b = np.asarray(a), np.asarray(b)
    sa = a.std(axis=0, ddof=1) / math.sqrt(len(a))
    sb = b.std(axis=0, ddof=1) / math.sqrt(len(b))
    return (a.mean(axis=0) - b.mean(axis=0)) / np.sqrt(sa ** 2 + sb ** 2)


def _device_pt(uc, L, S, temps, p, seed, precision=32):
    """one PT run of the device; per slot <E>, C_v and the exchange count"""
    lat = csmc.Lattice((L, L), uc, S, seed=seed)
    R = len(temps)
    b = csmc.SpinBatch([lat] * R, precision=precision)
    b.randomize(seed)
    b.pt_init(temps, p, csmc.METROPOLIS, seed=seed)
    while b.pt_advance(-1)[1] != 3:
        pass
    obs = b.pt_observables()
    st = b.pt_state()
    mode = b.storage_mode()
    b.close()
    return obs["energy"], obs["specific_heat"], st["accepted_exchange"], mode


def test_klein_duality_fp32_pt():
    L, S, R, n_runs = 16, 0.5, 10, 10
    temps = W.geometric_temperatures(0.05, 1.0, R)
    p = dict(t_thermalization=4000, t_measurement=40000, probe_rate=10, swap_rate=10,
             overrelaxation_rate=10, report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    kit = [_device_pt(W.honeycomb_kitaev(1.0, -2.0), L, S, temps, p, 100 + k) for k in range(n_runs)]
    fm = [_device_pt(W.honeycomb_kitaev(-1.0, 0.0), L, S, temps, p, 200 + k) for k in range(n_runs)]
    assert all(r[3] == 3 for r in kit + fm), "both models run in the register-table mode"
    for q, name in ((0, "<E>"), (1, "C_v"), (2, "exchange acceptance")):
        z = _z([r[q] for r in kit], [r[q] for r in fm])
        assert np.all(np.abs(z) < Z_MAX), (name, np.round(z, 2))
    # the duality also fixes the ground state, E/N = -(3/2) S^2 = -0.375, and
    # as T -> 0 each classical spin carries two quadratic modes, so
    # E/N -> -0.375 + T and C_v -> 1 (equipartition), with corrections of
    # order T / (J S^2) = 0.2 in the coldest slot
    N = 2 * L * L
    e_cold = np.mean([r[0][0] for r in kit]) / N
    c_cold = np.mean([r[1][0] for r in kit])
    assert -0.375 < e_cold < -0.375 + 1.5 * temps[0], e_cold
    assert 0.9 < c_cold < 1.2, c_cold


def _oracle_pt(uc, L, S, temps, t_th, t_meas, probe, rate, seed):
    """the reference's PT loop run by the oracle; the energy of every slot
    sampled every `probe` sweeps after t_th: per slot <E>, C_v and the
    exchange count"""
    T = orc.Tables(uc, (L, L))
    R = len(temps)
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(R, T.N, 3))
    spins = np.ascontiguousarray(S * v / np.linalg.norm(v, axis=2, keepdims=True))
    E = np.zeros(R)
    perm = np.arange(R, dtype=np.intc)
    acc_l = np.zeros(R, dtype=np.int_)
    acc_x = np.zeros(R, dtype=np.int_)
    series = []
    sw = 0
    while sw < t_th + t_meas:
        orc.pt_reference_run(T, spins, temps, S, sw, probe, rate, rate, seed, E, perm, acc_l, acc_x)
        sw += probe
        # sweep sw - probe + 1 .. sw - 1 ran; the probe of the reference is at
        # sweeps that are multiples of probe_rate, after the exchange
        if sw - probe >= t_th:
            series.append(E.copy())
    series = np.array(series)
    return series.mean(axis=0), series.var(axis=0) / (np.asarray(temps) ** 2 * T.N), acc_x


def test_cv_and_acceptance_vs_reference_chain():
    # honeycomb Kitaev-Heisenberg (the headline model) at L=4, S=1/2
    uc = W.honeycomb_kitaev(1.0, -2.0)
    L, S, R, n_runs = 4, 0.5, 6, 10
    temps = W.geometric_temperatures(0.08, 1.0, R)
    t_th, t_meas, rate = 2000, 40000, 10
    p = dict(t_thermalization=t_th, t_measurement=t_meas, probe_rate=rate, swap_rate=rate,
             overrelaxation_rate=rate, report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    dev = [_device_pt(uc, L, S, temps, p, 300 + k) for k in range(n_runs)]
    ref = [_oracle_pt(uc, L, S, temps, t_th, t_meas, rate, rate, 400 + k) for k in range(n_runs)]
    for q, name in ((0, "<E>"), (1, "C_v"), (2, "exchange acceptance")):
        z = _z([r[q] for r in dev], [r[q] for r in ref])
        assert np.all(np.abs(z) < Z_MAX), (name, np.round(z, 2))
