"""Checkpoint and initial-configuration files of parallel_tempering against
the oracle's replay of the reference's schedule.

The reference writes, once sweep >= t_thermalization and sweep %
checkpoint_rate == 0, after that sweep's exchange and before its probe
(monte_carlo.jl:353-365):
  - every rank's configuration into the `spins` dataset of its file
    (write_MC_checkpoint, hdf5.jl:176-180);
  - for the ranks in saveIC, IC_<rank>/IC_<k>.h5 with
    k = (sweep - t_thermalization) / checkpoint_rate, holding attributes T and
    paramsfile and the `spins` dataset (write_initial_configuration,
    hdf5.jl:185-194).
Each of those datasets must hold the replay's configuration of that
temperature slot at that sweep, and a run resumed from the file's saved
state must end where the replay ends."""
import os

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import hdf5io
from csmc import workloads as W
from conftest import random_spins
from pt_replay import colour_order, pt_replay

pytestmark = pytest.mark.gpu

TOL = 1e-10      # FP64 trajectories up to the last checkpoint (12 sweeps)
TOL_END = 1e-8   # the run's end, 16 sweeps: overrelaxation at these
                 # temperatures amplifies last-bit differences between the
                 # device's and the host's sin/cos about twofold per sweep


def _read_spins(path):
    with hdf5io.h5open(path, "r") as f:
        return np.asarray(f.read("spins")), f.get_attr("T")


def test_checkpoints_and_ics_equal_the_oracle_replay(tmp_path, monkeypatch):
    uc = W.kgg_test_cell()
    shape, S, R = (4, 4), 1.0, 5
    lat = csmc.Lattice(shape, uc, S, seed=2)
    T = orc.Tables(uc, shape)
    temps = W.geometric_temperatures(0.3, 3.0, R)
    params = dict(t_thermalization=4, t_measurement=12, probe_rate=2, swap_rate=2, overrelaxation_rate=2,
                  report_interval=10 ** 6, checkpoint_rate=4)
    seed = 31
    rng = np.random.default_rng(4)
    start = np.stack([random_spins(lat.size, S, rng) for _ in range(R)])
    out = str(tmp_path) + "/"
    save_ic = (0, 3)

    def make():
        mcs = []
        for r, t in enumerate(temps):
            m = csmc.MonteCarlo(t, lat, dict(params), outpath=out, rank=r, seed=seed, precision=64)
            m.lattice._spins[:] = start[r]
            mcs.append(m)
        return mcs

    # record every checkpoint write as the file holds it afterwards
    written = []
    real = hdf5io.write_MC_checkpoint

    def spy(mc):
        real(mc)
        written.append((mc.outpath, _read_spins(mc.outpath)[0].copy()))

    monkeypatch.setattr(hdf5io, "write_MC_checkpoint", spy)
    mcs = csmc.parallel_tempering(make(), saveIC=save_ic, verbose=False)
    total = params["t_thermalization"] + params["t_measurement"]
    cps = [s for s in range(total) if s >= params["t_thermalization"] and s % params["checkpoint_rate"] == 0]
    assert cps == [4, 8, 12]
    _, perm, E, _, _, snaps = pt_replay(T, start.copy(), colour_order(lat), temps, S, params, seed, 0,
                                        snapshots=set(cps))
    # the rank files: one spins write per rank and checkpoint, in sweep order
    assert len(written) == R * len(cps)
    for n, sw in enumerate(cps):
        for r in range(R):
            path, s = written[n * R + r]
            assert path == mcs[r].outpath
            assert np.max(np.abs(s - snaps[sw][r])) <= TOL * S, (sw, r)
    # the initial-configuration files of the saveIC ranks
    for r in save_ic:
        for sw in cps:
            k = (sw - params["t_thermalization"]) // params["checkpoint_rate"]
            path = os.path.join(out, f"IC_{r}", f"IC_{k}.h5")
            s, t_attr = _read_spins(path)
            assert t_attr == temps[r]
            assert np.max(np.abs(s - snaps[sw][r])) <= TOL * S, (r, sw)
        assert not os.path.exists(os.path.join(out, f"IC_{r}", f"IC_{len(cps)}.h5"))
    assert not os.path.exists(os.path.join(out, "IC_1"))
    monkeypatch.undo()
    # the run's end, and a fresh run resumed from the save at the last
    # checkpoint (sweep 12), both where the replay of all 16 sweeps ends
    _, perm_all, E_all, _, _, snaps_end = pt_replay(T, start.copy(), colour_order(lat), temps, S, params,
                                                    seed, 0, snapshots={total - 1})
    again = csmc.parallel_tempering(make(), verbose=False, resume=True)
    for r in range(R):
        assert np.max(np.abs(mcs[r].lattice._spins - snaps_end[total - 1][r])) <= TOL_END * S, r
        assert np.max(np.abs(again[r].lattice._spins - snaps_end[total - 1][r])) <= TOL_END * S, r
        assert np.array_equal(again[r].lattice._spins, mcs[r].lattice._spins), r
        assert again[r].pt_stats["system_of_slot"] == perm_all[r]
        assert mcs[r].pt_stats["system_of_slot"] == perm_all[r]
        assert abs(again[r].pt_stats["E"] - E_all[perm_all[r]]) <= TOL_END * abs(E_all[perm_all[r]])
