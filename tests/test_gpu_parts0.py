"""Chained parallel tempering with several workgroups per system when the
spins stay in device memory (storage mode 0): G workgroups share each
colour class of a system, a barrier among them ends every colour phase, and
the system's sums (accepted moves, FP64 energy, summed spin) are the parts'
workgroup sums added in part order.

The per-site updates and their Philox draws do not depend on the partition,
so the spins of every replica, the slot map and the exchange counts equal
those of the one-workgroup run bit for bit; the energies and observables are
summed in another order and agree to rounding."""
import os

import numpy as np
import pytest

import csmc
from csmc import workloads as W

pytestmark = pytest.mark.gpu


def _ctx(parts, extra=None):
    env = {"CSMC_LDS": "0", "CSMC_PARTS0": str(parts), **(extra or {})}
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


PARAMS = dict(t_thermalization=20, t_measurement=80, probe_rate=3, swap_rate=5, overrelaxation_rate=2,
              report_interval=10 ** 9, checkpoint_rate=10 ** 9)

# checkpoints every 7 sweeps end segments at host events: per-segment
# launches (with the exchange done by the last workgroup) between the chains
CHECKPOINTS = dict(PARAMS, checkpoint_rate=7)
NO_CHAIN = dict(PARAMS, checkpoint_rate=7)

CASES = [
    # (name, cell, shape, S, precision, replicas, parts, kind, parameters, environment)
    ("pyrochlore_fp32_R6_G4", W.pyrochlore_mixed, (3, 3, 3), 1.0, 32, 6, 4, csmc.METROPOLIS, PARAMS, {}),
    ("pyrochlore_fp64_R4_G3", W.pyrochlore_mixed, (2, 2, 2), 1.0, 64, 4, 3, csmc.CONE, PARAMS, {}),
    ("honeycomb_fp32_R8_G2", lambda: W.honeycomb_kitaev(1.0, -2.0), (24, 24), 0.5, 32, 8, 2,
     csmc.CONE_ADAPTIVE, PARAMS, {}),
    ("square_fp64_R5_G8", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (40, 40), 1.0, 64, 5, 8,
     csmc.METROPOLIS, PARAMS, {}),
    ("pyrochlore_fp32_R4_G4_checkpoints", W.pyrochlore_mixed, (3, 3, 3), 1.0, 32, 4, 4, csmc.METROPOLIS,
     CHECKPOINTS, {}),
    ("honeycomb_fp32_R3_G3_no_chain", lambda: W.honeycomb_kitaev(1.0, -2.0), (24, 24), 0.5, 32, 3, 3,
     csmc.METROPOLIS, NO_CHAIN, {"CSMC_PT_CHAIN": "0"}),
]


def _rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b)))))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_parts_equal_one_workgroup_per_system(case):
    name, mk, shape, S, prec, R, G, kind, params, extra = case
    print(f"[parts0] {name}", flush=True)
    lat = csmc.Lattice(shape, mk(), S, seed=4)
    temps = W.geometric_temperatures(1.0, 1.5, R)
    runs = []
    for parts in (1, G):
        ctx = _ctx(parts, extra)
        b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
        b.randomize(21)
        b.pt_init(temps, params, kind, seed=8)
        assert b.pt_split_parts() == parts
        events = []
        while True:
            stopped, ev = b.pt_advance(-1)
            events.append(ev)
            if ev == 3:
                break
        assert (1 in events) == (params is not PARAMS)
        runs.append(dict(spins=[b.get_spins(r) for r in range(R)], state=b.pt_state(),
                         obs=b.pt_observables(), E=b.total_energy()))
        b.close()
        ctx.close()
    one, many = runs
    for r in range(R):
        assert np.array_equal(one["spins"][r], many["spins"][r]), (name, r)
    for key in ("system_of_slot", "accepted_exchange"):
        assert np.array_equal(one["state"][key], many["state"][key]), key
    assert one["state"]["accepted_exchange"].sum() > 0
    assert _rel(one["state"]["accepted_local"], many["state"]["accepted_local"]) == 0.0
    assert np.array_equal(one["E"], many["E"])
    for key in one["obs"]:
        assert _rel(many["obs"][key], one["obs"][key]) <= 1e-9, key


SWEEP_CASES = [
    # (name, cells and shapes, precision, parts)
    ("pyrochlore_honeycomb_fp32_G4", [(W.pyrochlore_mixed, (3, 3, 3), 1.0),
                                      (lambda: W.honeycomb_kitaev(1.0, -2.0), (20, 20), 0.5)], 32, 4),
    ("square_fp64_G3", [(lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (36, 36), 1.0)], 64, 3),
]


@pytest.mark.parametrize("case", SWEEP_CASES, ids=[c[0] for c in SWEEP_CASES])
def test_parts_sweeps_and_annealing_equal_one_workgroup(case):
    """Sweeps of every kind and a whole annealing schedule in one launch, with
    the same parts: spins and accepted counts equal to one workgroup per
    system, bit for bit"""
    name, cells, prec, G = case
    print(f"[parts0] sweeps {name}", flush=True)
    lats = [csmc.Lattice(shape, mk(), S, seed=6 + k) for k, (mk, shape, S) in enumerate(cells)]
    T = np.linspace(0.4, 1.1, len(lats))
    runs = []
    for parts in (1, G):
        ctx = _ctx(parts)
        b = csmc.SpinBatch(lats, precision=prec, ctx=ctx)
        b.randomize(13)
        accs = []
        for kind in (csmc.OVERRELAX, csmc.METROPOLIS, csmc.CONE, csmc.DETERMINISTIC):
            accs.append(b.sweep(kind, T=T, sigma=0.7, n_sweeps=3, seed=19))
        sg = np.full(len(lats), 0.9)
        accs.append(b.sweep(csmc.CONE_ADAPTIVE, T=T, sigma=sg, n_sweeps=4, seed=19))
        accs.append(sg.copy())
        accs.append(b.anneal(W.geometric_temperatures(0.2, 1.0, 4)[::-1], 6, 2, csmc.METROPOLIS, seed=3))
        runs.append(([b.get_spins(r) for r in range(len(lats))], accs))
        b.close()
        ctx.close()
    (s1, a1), (sg_, ag) = runs
    for r in range(len(lats)):
        assert np.array_equal(s1[r], sg_[r]), (name, r)
    for x, y in zip(a1, ag):
        assert np.array_equal(x, y)
    assert a1[1].sum() > 0 and a1[-1].sum() > 0
