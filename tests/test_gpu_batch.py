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


def _ctx(env):
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return csmc.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("precision", [32, 64])
def test_packed_schedule_equals_one_workgroup_per_system(precision):
    # a batch of one large lattice (L = 128, 16384 sites: storage mode 0) and
    # many small ones: the packed schedule splits the large lattice into parts
    # on one XCD and gives the small ones to other workgroups.  Sweeps and a
    # whole annealing schedule must equal, bit for bit, the schedule with one
    # workgroup per system (CSMC_PACK=0): each system's draws and updates do
    # not depend on where, or in how many parts, it runs.
    cells = [(W.square_dm(-1.0, 0.05, [0.1, 0.0, 0.2]), (128, 128))]
    cells += W.uneven_batch(n=40, sizes=(16, 24, 32), seed=9)
    lats = [csmc.Lattice(shape, uc, 1.0, seed=30 + k) for k, (uc, shape) in enumerate(cells)]
    out = []
    for env in ({}, {"CSMC_PACK": "0"}):
        print(f"[packed] precision={precision} env={env}", flush=True)
        ctx = _ctx(dict(env, CSMC_VERBOSE="1"))
        b = csmc.SpinBatch(lats, precision=precision, ctx=ctx)
        b.randomize(11)
        assert b.storage_mode() == 0
        b.sweep(csmc.OVERRELAX, n_sweeps=3)
        acc_s = b.sweep(csmc.METROPOLIS, T=np.linspace(0.2, 2.0, len(lats)), n_sweeps=2, seed=5)
        acc_a = b.anneal([1.0, 0.6, 0.3], 12, 3, csmc.METROPOLIS, seed=7)
        out.append(([b.get_spins(k) for k in range(len(lats))], b.counters(), acc_s, acc_a))
        b.close()
        ctx.close()
    (s0, c0, as0, aa0), (s1, c1, as1, aa1) = out
    assert np.array_equal(c0, c1) and np.array_equal(as0, as1) and np.array_equal(aa0, aa1)
    for k in range(len(lats)):
        assert np.array_equal(s0[k], s1[k]), k
