"""Spin correlations inside parallel tempering (MonteCarlo(...; corr = true,
ks), monte_carlo.jl:371-375; output hdf5.jl:229-233) and the runners over
saved configurations (spin_correlations.jl:48-145).

The device sums S(k) of every system into the temperature slot it holds at
each probe sweep, after that sweep's exchange.  The replay steps the same run
one sweep at a time, reads the slot map and the spins after every probe
sweep, and averages the oracle's correlations of those spins per slot."""
import os

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import hdf5io as H
from csmc import workloads as W
from csmc.reciprocal import get_allowed_wavevectors

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=1e-12):
    return np.max(np.abs(a - b)) <= rtol * max(1.0, float(np.max(np.abs(b))))


PARAMS = dict(t_thermalization=20, t_measurement=60, probe_rate=5, swap_rate=10, overrelaxation_rate=3,
              report_interval=10 ** 9, checkpoint_rate=10 ** 9)


def _ctx(split):
    old = os.environ.get("CSMC_SPLIT")
    os.environ["CSMC_SPLIT"] = split
    try:
        return csmc.Context(0)
    finally:
        if old is None:
            os.environ.pop("CSMC_SPLIT", None)
        else:
            os.environ["CSMC_SPLIT"] = old


# probes sparser than exchanges, so that chains of exchange periods (and, with
# CSMC_SPLIT=4, the split kernel) run between them
SPARSE = dict(PARAMS, probe_rate=20, swap_rate=5, t_measurement=100)


@pytest.mark.parametrize("prec,shape,split,params", [(64, (6, 6), "1", PARAMS), (32, (6, 6), "1", PARAMS),
                                                     (32, (32, 8), "4", SPARSE)],
                         ids=["fp64", "fp32", "fp32_split4_chains"])
def test_pt_correlations_match_a_stepped_oracle_replay(prec, shape, split, params):
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice(shape, uc, 0.5, seed=3)
    R = 6
    temps = W.geometric_temperatures(0.2, 2.0, R)
    ks = get_allowed_wavevectors(uc, shape)
    total = params["t_thermalization"] + params["t_measurement"]
    print(f"[pt_corr] stepped replay fp{prec} {shape} split {split}", flush=True)
    ctx = _ctx(split)
    b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
    b.randomize(17)
    b.pt_init(temps, params, csmc.METROPOLIS, seed=23)
    b.pt_set_correlations(ks)
    acc = np.zeros((R, 9, ks.shape[1]))
    n = 0
    for s in range(total):
        stopped, ev = b.pt_advance(s + 1)
        assert stopped == s + 1
        if s >= params["t_thermalization"] and s % params["probe_rate"] == 0:
            sos = b.pt_state()["system_of_slot"]
            for T in range(R):
                acc[T] += orc.equal_time_correlations(lat.site_positions, b.get_spins(int(sos[T])), ks)
            n += 1
    mean, count = b.pt_correlations()
    assert count == n == (12 if params is PARAMS else 5)
    assert mean.shape == (R, 9, ks.shape[1])
    assert _close(mean, acc / n)
    assert b.pt_state()["accepted_exchange"].sum() > 0
    if split != "1":
        assert b.pt_split_parts() > 1
    b.close()

    # the same run in one call (chained launches between the probes)
    b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
    b.randomize(17)
    b.pt_init(temps, params, csmc.METROPOLIS, seed=23)
    b.pt_set_correlations(ks)
    stopped, ev = b.pt_advance(-1)
    assert ev == 3
    mean2, count2 = b.pt_correlations()
    assert count2 == n
    if split == "1":
        assert np.array_equal(mean2, mean)
    else:
        # the split kernel sums the energies the exchanges use in another
        # order: the same trajectory here, its S(k) equal to rounding
        assert _close(mean2, mean)
    b.close()
    ctx.close()


def test_correlations_off_and_state_errors():
    lat = csmc.Lattice((4, 4), W.square_heisenberg(-1.0, (0, 0, 0.1)), 1.0, seed=1)
    b = csmc.SpinBatch([lat] * 2, precision=64)
    b.randomize(1)
    with pytest.raises(csmc.CsmcError, match="csmc_pt_init"):
        b.pt_set_correlations(np.zeros((2, 3)))
    b.pt_init(np.array([0.5, 1.0]), PARAMS, csmc.METROPOLIS, seed=2)
    with pytest.raises(csmc.CsmcError, match="not requested"):
        b.pt_correlations()
    b.pt_advance(5)
    with pytest.raises(csmc.CsmcError, match="before the first"):
        b.pt_set_correlations(np.zeros((2, 3)))
    b.close()


def test_parallel_tempering_corr_writes_ssf_and_runners(tmp_path):
    """parallel_tempering with corr = true writes each slot's mean S(k) to
    spin_correlations/SSF; the saved configurations of slot 0 then go through
    runEqualTimeStructureFactor and compute_equal_time_structure_factor"""
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice((6, 6), uc, 0.5, seed=5)
    ks = get_allowed_wavevectors(uc, (6, 6))
    temps = W.geometric_temperatures(0.3, 1.5, 4)
    out = str(tmp_path) + "/"
    params = dict(t_thermalization=20, t_measurement=40, probe_rate=4, swap_rate=5, overrelaxation_rate=2,
                  report_interval=10 ** 6, checkpoint_rate=10)
    mcs = [csmc.MonteCarlo(T, lat, dict(params), outpath=out, rank=r, corr=True, ks=ks, seed=3)
           for r, T in enumerate(temps)]
    print("[pt_corr] parallel_tempering with corr and saveIC", flush=True)
    mcs = csmc.parallel_tempering(mcs, saveIC=(0,), verbose=False)
    for m in mcs:
        assert m.correlations.shape == (9, ks.shape[1])
        assert m.correlation_count == 10  # sweeps 20, 24, ..., 56
        with H.h5open(m.outpath, "r") as f:
            assert np.array_equal(f.read("spin_correlations/SSF"), m.correlations.T)
            assert np.array_equal(f.read("spin_correlations/SSF_momentum"), ks.T)
        # S^zz(0) of a configuration of N spins of length S lies in [0, N S^2]
        assert np.all(np.isfinite(m.correlations))
    icdir = os.path.join(str(tmp_path), "IC_0") + "/"
    files = sorted(os.listdir(icdir))
    assert len(files) == 4  # checkpoints at sweeps 20, 30, 40, 50
    csmc.runEqualTimeStructureFactor(icdir, csmc.Lattice((6, 6), uc, 0.5, seed=5), ks)
    refs = []
    for name in files:
        with H.h5open(icdir + name, "r") as f:
            spins = f.read("spins")
            got = f.read("spin_correlations/SSF").T
        ref = orc.equal_time_correlations(lat.site_positions, spins, ks)
        assert _close(got, ref), name
        refs.append(ref)
    # a second run skips every file
    csmc.runEqualTimeStructureFactor(icdir, csmc.Lattice((6, 6), uc, 0.5, seed=5), ks)
    mean, kout = csmc.compute_equal_time_structure_factor(icdir, mcs[0].outpath)
    assert _close(mean, sum(refs) / len(refs))
    assert np.array_equal(kout, ks)
