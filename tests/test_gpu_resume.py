"""Resume of a parallel-tempering run (csmc_pt_save / _save_async /
_save_wait / _load; no counterpart in the reference, whose checkpoints hold
the spins only, hdf5.jl:176-180).

A run stopped part-way, saved, and restored into a fresh batch continues
exactly as the uninterrupted run: spins, counters, slot maps, energies and
correlation sums bit for bit, the binned observables to 1e-12 (their |M|
can come from a chained or a single-period launch, which the compiler may
contract differently).  A save taken asynchronously, with later launches
queued behind its copy, equals a synchronous save at the same point
byte for byte.  Saved states that do not match the run are refused."""
import os

import numpy as np
import pytest

import csmc
from csmc import workloads as W
from csmc.reciprocal import get_allowed_wavevectors

pytestmark = pytest.mark.gpu


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


PARAMS = dict(t_thermalization=30, t_measurement=120, probe_rate=5, swap_rate=10, overrelaxation_rate=3,
              report_interval=10 ** 9, checkpoint_rate=10 ** 9)

CASES = [  # name, shape, precision, replicas, CSMC_SPLIT, stop sweep, correlations
    ("fp32_R8_whole", (16, 16), 32, 8, "1", 61, True),
    ("fp64_R5_whole", (8, 8), 64, 5, "1", 47, False),
    ("fp32_R8_split4", (32, 8), 32, 8, "4", 75, False),
]


def _start(ctx, lat, R, prec, corr, ks, temps):
    b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
    b.randomize(17)
    b.pt_init(temps, PARAMS, csmc.METROPOLIS, seed=23)
    if corr:
        b.pt_set_correlations(ks)
    return b


def _result(b, R, corr):
    st = b.pt_state()
    out = {"state": st, "obs": b.pt_observables(), "spins": [b.get_spins(r) for r in range(R)],
           "counters": b.counters()}
    if corr:
        out["corr"] = b.pt_correlations()
    return out


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_saved_run_continues_as_the_uninterrupted_one(case):
    name, shape, prec, R, split, stop, corr = case
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice(shape, uc, 0.5, seed=3)
    ks = get_allowed_wavevectors(uc, shape)
    temps = W.geometric_temperatures(0.2, 2.0, R)
    print(f"[resume] {name}", flush=True)
    ctx = _ctx(split)
    # uninterrupted
    a = _start(ctx, lat, R, prec, corr, ks, temps)
    assert a.pt_advance(-1)[1] == 3
    ref = _result(a, R, corr)
    a.close()
    # stopped at `stop`, saved both ways, then run on
    b = _start(ctx, lat, R, prec, corr, ks, temps)
    b.pt_advance(stop)
    sync = b.pt_save()
    b.pt_save_async()
    b.pt_advance(stop + 20)  # later launches queued behind the copy
    assert b.pt_save_wait() == sync
    b.close()
    # restored into a fresh batch (and a fresh context) and run to the end
    ctx2 = _ctx(split)
    c = _start(ctx2, lat, R, prec, corr, ks, temps)
    c.pt_load(sync)
    stopped, ev = c.pt_advance(-1)
    assert ev == 3
    got = _result(c, R, corr)
    c.close()
    for k in ref["state"]:
        assert np.array_equal(got["state"][k], ref["state"][k]), k
    for k in ref["obs"]:
        assert np.allclose(got["obs"][k], ref["obs"][k], rtol=1e-12, atol=0.0), k
    for x, y in zip(got["spins"], ref["spins"]):
        assert np.array_equal(x, y)
    assert np.array_equal(got["counters"], ref["counters"])
    if corr:
        assert got["corr"][1] == ref["corr"][1]
        assert np.array_equal(got["corr"][0], ref["corr"][0])
    assert ref["state"]["accepted_exchange"].sum() > 0
    ctx.close()
    ctx2.close()


def test_mismatched_saved_state_is_refused():
    lat = csmc.Lattice((8, 8), W.square_heisenberg(-1.0, (0, 0, 0.1)), 1.0, seed=1)
    temps = W.geometric_temperatures(0.3, 1.0, 4)
    b = csmc.SpinBatch([lat] * 4, precision=64)
    b.randomize(1)
    b.pt_init(temps, PARAMS, csmc.METROPOLIS, seed=5)
    b.pt_advance(12)
    state = b.pt_save()
    b.close()

    def fresh(prec=64, seed=5, t=temps, n=4, params=PARAMS):
        c = csmc.SpinBatch([lat] * n, precision=prec)
        c.pt_init(t, params, csmc.METROPOLIS, seed=seed)
        return c

    for kw, msg in [(dict(seed=6), "seed"), (dict(t=temps * 1.01), "temperatures"),
                    (dict(prec=32), "precision"), (dict(n=3, t=temps[:3]), "number of systems"),
                    (dict(params=dict(PARAMS, swap_rate=5)), "parameters")]:
        c = fresh(**kw)
        with pytest.raises(csmc.CsmcError, match=msg):
            c.pt_load(state)
        c.close()
    c = fresh()
    with pytest.raises(csmc.CsmcError, match="not a saved"):
        c.pt_load(b"\0" * 200)
    with pytest.raises(csmc.CsmcError, match="no save in flight"):
        c.pt_save_wait()
    c.pt_load(state)  # the matching run accepts it
    c.close()


def test_parallel_tempering_resumes_from_its_checkpoint_file(tmp_path):
    """the driver saves the run at every checkpoint; fresh MonteCarlo objects
    on the same files with resume=True continue from the last save and end
    where the uninterrupted run ended"""
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice((8, 8), uc, 0.5, seed=5)
    temps = W.geometric_temperatures(0.3, 1.5, 4)
    out = str(tmp_path) + "/"
    params = dict(t_thermalization=20, t_measurement=70, probe_rate=5, swap_rate=5, overrelaxation_rate=2,
                  report_interval=10 ** 6, checkpoint_rate=25)

    def mcs():
        return [csmc.MonteCarlo(T, lat, dict(params), outpath=out, rank=r, seed=3, precision=32)
                for r, T in enumerate(temps)]

    print("[resume] driver run with checkpoints", flush=True)
    first = csmc.parallel_tempering(mcs(), verbose=False)
    saved = csmc.hdf5io.read_resume_state(first[0])
    assert saved is not None and saved[1] == 75  # checkpoints at 25, 50, 75
    print("[resume] driver resumed from the file", flush=True)
    again = csmc.parallel_tempering(mcs(), verbose=False, resume=True)
    for x, y in zip(first, again):
        assert np.array_equal(x.lattice._spins, y.lattice._spins)
        for k in x.observables:
            assert np.isclose(x.observables[k], y.observables[k], rtol=1e-12, atol=0.0), k
