"""The storage modes of the persistent sweep kernel give the same FP32
results: the register-table mode (neighbour offsets in registers, spins as
float4 in LDS, diagonal couplings specialised) against the LDS-array mode on
the same inputs, bit for bit for sweeps and the parallel-tempering schedule
(the FP64 energies, summed in a different order, agree to 1e-12)."""
import os

import numpy as np
import pytest

import csmc
from csmc import workloads as W

pytestmark = pytest.mark.gpu


def _ctx(cap):
    old = os.environ.get("CSMC_LDS")
    os.environ["CSMC_LDS"] = str(cap)
    try:
        return csmc.Context(0)
    finally:
        if old is None:
            del os.environ["CSMC_LDS"]
        else:
            os.environ["CSMC_LDS"] = old


CASES = [  # (name, unit cell factory, shape, S, expected modes (cap 3, cap 2))
    ("honeycomb_kitaev_L64", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 64), 0.5, (3, 2)),
    ("kgg_L8", W.kgg_test_cell, (8, 8), 1.0, (3, 2)),
    ("square_L16", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (16, 16), 1.0, (3, 2)),
    ("triangular_dm_L24", lambda: W.triangular_dm(-0.9, -0.05, [0.3, 0.1, 0.0]), (24, 24), 1.0, (3, 2)),
    ("square_L96_fallback", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (96, 96), 1.0, (1, 1)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_register_mode_matches_lds_mode(case):
    name, mk, shape, S, modes = case
    lat = csmc.Lattice(shape, mk(), S, seed=1)
    R = 4
    out = []
    for cap, mode in zip((3, 2), modes):
        ctx = _ctx(cap)
        b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
        assert b.storage_mode() == mode
        b.randomize(11)
        b.set_counters(0)
        T = np.array([0.2, 0.5, 1.0, 2.0])
        acc = [b.sweep(csmc.OVERRELAX, n_sweeps=3)]
        acc.append(b.sweep(csmc.METROPOLIS, T=T, n_sweeps=2, seed=5))
        sg = np.full(R, 0.3)
        acc.append(b.sweep(csmc.CONE_ADAPTIVE, T=T, sigma=sg, n_sweeps=2, seed=6))
        acc.append(b.sweep(csmc.DETERMINISTIC, n_sweeps=1))
        acc.append(b.anneal([0.8, 0.4], 5, 2, csmc.METROPOLIS, seed=7).ravel())
        spins = [b.get_spins(r) for r in range(R)]
        p = dict(t_thermalization=6, t_measurement=20, probe_rate=2, swap_rate=4, overrelaxation_rate=2,
                 report_interval=10 ** 9, checkpoint_rate=10 ** 9)
        b.pt_init(W.geometric_temperatures(0.3, 2.0, R), p, csmc.METROPOLIS, seed=9)
        b.pt_advance(-1)
        st = b.pt_state()
        out.append((acc, sg.copy(), spins, st, [b.get_spins(r) for r in range(R)]))
        b.close()
        ctx.close()
    (a3, s3, sp3, st3, pt3), (a2, s2, sp2, st2, pt2) = out
    for x, y in zip(a3, a2):
        assert np.array_equal(x, y)
    assert np.array_equal(s3, s2)
    for x, y in zip(sp3, sp2):
        assert np.array_equal(x, y)
    assert np.array_equal(st3["system_of_slot"], st2["system_of_slot"])
    assert np.array_equal(st3["accepted_local"], st2["accepted_local"])
    assert np.array_equal(st3["accepted_exchange"], st2["accepted_exchange"])
    assert np.allclose(st3["E"], st2["E"], rtol=1e-12, atol=1e-12)
    for x, y in zip(pt3, pt2):
        assert np.array_equal(x, y)
