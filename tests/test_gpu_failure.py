"""Failure path of the waits between workgroups of a chained PT launch: with
the wait limit forced to zero from the host (CSMC_SPIN_LIMIT=0), the first
wait that does not find its partner's record at once gives up, sets the error
flag, and every later wait sees the flag and returns -- so every wave still
reaches the end of the kernel -- and csmc_pt_advance reports the failure.  A
fresh context on the same device then runs normally.  Nothing faults: the
waits are bounded by design and the limit only shortens them."""
import os

import numpy as np
import pytest

import csmc
from csmc import _lib, workloads as W

pytestmark = pytest.mark.gpu


def _ctx(env):
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


@pytest.mark.parametrize("split", ["1", "4"], ids=["whole_system", "split_G4"])
def test_wait_timeout_is_reported_and_device_stays_usable(split):
    lat = csmc.Lattice((32, 8), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=3)
    R = 16
    temps = W.geometric_temperatures(0.1, 2.0, R)
    p = dict(t_thermalization=10, t_measurement=2000, probe_rate=10, swap_rate=10, overrelaxation_rate=2,
             report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    print(f"[failure] split={split}", flush=True)
    ctx = _ctx({"CSMC_SPIN_LIMIT": "0", "CSMC_SPLIT": split})
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(5)
    b.pt_init(temps, p, csmc.METROPOLIS, seed=7)
    with pytest.raises(_lib.CsmcError, match="timed out"):
        b.pt_advance(-1)
    b.close()
    ctx.close()
    # the device is fine: the same run on a fresh context completes
    ctx = _ctx({"CSMC_SPLIT": split})
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(5)
    b.pt_init(temps, p, csmc.METROPOLIS, seed=7)
    stopped, ev = b.pt_advance(-1)
    assert ev == 3
    st = b.pt_state()
    assert np.all(np.isfinite(st["E"])) and st["accepted_exchange"].sum() > 0
    b.close()
    ctx.close()


def test_parts_wait_timeout_in_sweeps_and_annealing():
    # storage mode 0 with four workgroups per system (CSMC_LDS=0,
    # CSMC_PARTS0=4): the barrier between the parts of a system ends every
    # colour phase; with the wait limit forced to zero the first wait that
    # does not see the phase complete gives up and sets the batch's flag, and
    # csmc_sweep and csmc_anneal report it.  A fresh context then runs the
    # same calls normally.
    lat = csmc.Lattice((32, 16), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=3)
    R = 4
    print("[failure] parts in sweeps and annealing", flush=True)
    ctx = _ctx({"CSMC_SPIN_LIMIT": "0", "CSMC_LDS": "0", "CSMC_PARTS0": "4"})
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(5)
    assert b.storage_mode() == 0
    with pytest.raises(_lib.CsmcError, match="timed out"):
        b.sweep(csmc.OVERRELAX, n_sweeps=4)
    with pytest.raises(_lib.CsmcError, match="timed out"):
        b.anneal([1.0, 0.5], 20, 2, csmc.METROPOLIS, seed=3)
    b.close()
    ctx.close()
    ctx = _ctx({"CSMC_LDS": "0", "CSMC_PARTS0": "4"})
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(5)
    b.sweep(csmc.OVERRELAX, n_sweeps=4)
    acc = b.anneal([1.0, 0.5], 20, 2, csmc.METROPOLIS, seed=3)
    assert np.all(acc > 0)
    assert np.allclose(np.linalg.norm(b.get_spins(0), axis=1), 0.5, atol=1e-5)
    b.close()
    ctx.close()


def test_parts_chosen_again_when_the_storage_mode_drops_to_global():
    # a batch whose PT state was set up in an LDS storage mode (no parts) and
    # is then autotuned: its storage-mode-0 trial must run with the parts
    # CSMC_PARTS0 asks for, so with the wait limit forced to zero that trial
    # reports the timed-out wait (a trial without parts has no wait to time out)
    lat = csmc.Lattice((32, 16), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=3)
    R = 4
    temps = W.geometric_temperatures(0.2, 1.0, R)
    p = dict(t_thermalization=10, t_measurement=10, probe_rate=5, swap_rate=5, overrelaxation_rate=2,
             report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    print("[failure] parts after an LDS-mode PT setup", flush=True)
    ctx = _ctx({"CSMC_SPIN_LIMIT": "0", "CSMC_PARTS0": "4"})
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(8)
    b.pt_init(temps, p, csmc.METROPOLIS, seed=2)
    assert b.storage_mode() > 0 and b.pt_split_parts() == 1
    with pytest.raises(_lib.CsmcError, match="timed out"):
        b.autotune(csmc.OVERRELAX, n_sweeps=2)
    b.close()
    ctx.close()
