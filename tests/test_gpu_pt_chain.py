"""Chained parallel tempering (one cooperative launch for many exchange
periods, each exchange decided pair by pair from records the partner slots
publish) against one launch per period (the last workgroup decides every
pair): the same schedule -- spins, counters, slot maps and energies bit for
bit, the binned observables to 1e-12."""
import os

import numpy as np
import pytest

import csmc
from csmc import workloads as W

pytestmark = pytest.mark.gpu


def _ctx(chain):
    # one workgroup per system in both forms (the split form sums energies
    # over its parts in another order; it has its own test)
    old = {k: os.environ.get(k) for k in ("CSMC_PT_CHAIN", "CSMC_SPLIT")}
    os.environ["CSMC_PT_CHAIN"] = "1" if chain else "0"
    os.environ["CSMC_SPLIT"] = "1"
    try:
        return csmc.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


CASES = [  # name, cell, shape, S, precision, replicas, alg, schedule
    ("honeycomb_fp32_R16", lambda: W.honeycomb_kitaev(1.0, -2.0), (16, 16), 0.5, 32, 16, csmc.METROPOLIS,
     dict(t_thermalization=40, t_measurement=200, probe_rate=10, swap_rate=10, overrelaxation_rate=10,
          report_interval=10 ** 9, checkpoint_rate=10 ** 9)),
    ("square_fp64_R7_events", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (8, 8), 1.0, 64, 7,
     csmc.CONE_ADAPTIVE,
     dict(t_thermalization=13, t_measurement=120, probe_rate=3, swap_rate=4, overrelaxation_rate=2,
          report_interval=50, checkpoint_rate=25)),
    ("pyrochlore_fp32_R5", W.pyrochlore_mixed, (2, 2, 2), 1.0, 32, 5, csmc.METROPOLIS,
     dict(t_thermalization=10, t_measurement=60, probe_rate=2, swap_rate=3, overrelaxation_rate=1,
          report_interval=10 ** 9, checkpoint_rate=10 ** 9)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_chained_pt_matches_per_period_launches(case):
    name, mk, shape, S, prec, R, alg, p = case
    lat = csmc.Lattice(shape, mk(), S, seed=3)
    temps = W.geometric_temperatures(0.2, 3.0, R)
    out = []
    for chain in (True, False):
        ctx = _ctx(chain)
        b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
        b.randomize(17)
        b.pt_init(temps, p, alg, sigma0=0.6, seed=23)
        events = []
        while True:
            stopped, ev = b.pt_advance(-1)
            events.append((stopped, ev))
            if ev == 3:
                break
        out.append((events, b.pt_state(), b.pt_observables(), [b.get_spins(r) for r in range(R)],
                    b.counters()))
        b.close()
        ctx.close()
    (e1, st1, ob1, sp1, c1), (e2, st2, ob2, sp2, c2) = out
    assert e1 == e2
    for k in st1:
        assert np.array_equal(st1[k], st2[k]), k
    # the binned observables come out of two different kernels, whose FP64
    # |M| = sqrt(mx^2 + my^2 + mz^2) the compiler may contract differently:
    # equal to the last few bits
    for k in ob1:
        assert np.allclose(ob1[k], ob2[k], rtol=1e-12, atol=0.0), k
    for x, y in zip(sp1, sp2):
        assert np.array_equal(x, y)
    assert np.array_equal(c1, c2)
    assert st1["accepted_exchange"].sum() > 0
