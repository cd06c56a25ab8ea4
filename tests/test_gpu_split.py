"""The split form of the chained parallel-tempering kernel (each system
shared by G workgroups, strips with halos, strip edges and partial sums
swapped through epoch-tagged words at every window end) against the
whole-system kernel on the same inputs: spins, Metropolis counters, slot
maps, accepted counts and sigmas bit for bit; the FP64 energies, summed
over the parts in another order, to 1e-12 (the binned observables derived
from them to 1e-7)."""
import os

import numpy as np
import pytest

import csmc
from csmc import workloads as W
from csmc.unitcell import Pyrochlore, addBilinear

pytestmark = pytest.mark.gpu


def _ctx(split):
    old = os.environ.get("CSMC_SPLIT")
    if split is None:
        os.environ.pop("CSMC_SPLIT", None)
    else:
        os.environ["CSMC_SPLIT"] = str(split)
    try:
        return csmc.Context(0)
    finally:
        if old is None:
            os.environ.pop("CSMC_SPLIT", None)
        else:
            os.environ["CSMC_SPLIT"] = old


def pyrochlore_heisenberg():
    uc = Pyrochlore()
    for b1, b2, o in W.PYRO_BONDS:
        addBilinear(uc, b1, b2, np.eye(3), o)
    return uc


def sched(t_th, t_meas, probe, swap, orr, report=10 ** 9, ckpt=10 ** 9):
    return dict(t_thermalization=t_th, t_measurement=t_meas, probe_rate=probe, swap_rate=swap,
                overrelaxation_rate=orr, report_interval=report, checkpoint_rate=ckpt)


CASES = [  # name, cell, shape, S, replicas, CSMC_SPLIT (None: unset), expected parts, alg, schedule
    ("honeycomb_64x8_G4_aligned", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 8), 0.5, 8, 4, 4,
     csmc.METROPOLIS, sched(40, 400, 10, 10, 10)),
    ("honeycomb_32x8_G3_unaligned_probes_adaptive", lambda: W.honeycomb_kitaev(1.0, -2.0, 0.3, -0.05),
     (32, 8), 0.5, 6, 3, 3, csmc.CONE_ADAPTIVE, sched(12, 150, 3, 4, 2)),
    ("square_24x6_G2_no_overrelaxation", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (24, 6), 1.0,
     5, 2, 2, csmc.METROPOLIS, sched(5, 60, 1, 2, 0)),
    ("triangular_18x6_G3_events", lambda: W.triangular_dm(-0.9, -0.05, [0.3, 0.1, 0.0]), (18, 6), 1.0,
     4, 3, 3, csmc.CONE, sched(10, 110, 2, 3, 1, report=37, ckpt=50)),
    ("pyrochlore_8x3x3_G2", pyrochlore_heisenberg, (8, 3, 3), 1.0, 4, 2, 2, csmc.METROPOLIS,
     sched(6, 80, 2, 4, 2)),
    ("honeycomb_64x64_R64_auto", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 64), 0.5, 64, "auto", 4,
     csmc.METROPOLIS, sched(20, 200, 10, 10, 10)),
    ("honeycomb_64x64_R64_default_whole", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 64), 0.5, 64, None, 1,
     csmc.METROPOLIS, sched(10, 60, 10, 10, 10)),
    ("honeycomb_64x64_R64_G8_not_resident", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 64), 0.5, 64, 8,
     1, csmc.METROPOLIS, sched(10, 60, 10, 10, 10)),
]


def run(lat, R, prec, split, alg, p):
    ctx = _ctx(split)
    b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
    b.randomize(17)
    b.pt_init(W.geometric_temperatures(0.05, 2.0, R), p, alg, sigma0=0.6, seed=23)
    parts0 = b.pt_split_parts()
    events = []
    while True:
        stopped, ev = b.pt_advance(-1)
        events.append((stopped, ev))
        if ev == 3:
            break
    out = dict(events=events, parts0=parts0, parts=b.pt_split_parts(), state=b.pt_state(),
               obs=b.pt_observables(), spins=[b.get_spins(r) for r in range(R)], counters=b.counters(),
               energy=b.total_energy())
    b.close()
    ctx.close()
    return out


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_split_matches_whole_system_kernel(case):
    name, mk, shape, S, R, split, parts, alg, p = case
    lat = csmc.Lattice(shape, mk(), S, seed=3)
    print(f"[split] {name}", flush=True)
    whole = run(lat, R, 32, 1, alg, p)
    part = run(lat, R, 32, split, alg, p)
    assert whole["parts0"] == 1
    assert part["parts"] == parts, "split kernel not used"
    if isinstance(split, int) and parts == 1:
        assert part["parts0"] == split  # planned, refused at launch, fell back
    assert part["events"] == whole["events"]
    for x, y in zip(part["spins"], whole["spins"]):
        assert np.array_equal(x, y)
    assert np.array_equal(part["counters"], whole["counters"])
    for k in ("system_of_slot", "accepted_local", "accepted_exchange", "sigma"):
        assert np.array_equal(part["state"][k], whole["state"][k]), k
    E0 = whole["state"]["E"]
    assert np.allclose(part["state"]["E"], E0, rtol=1e-12, atol=1e-12 * np.abs(E0).max())
    # the observables are variances (and errors of variances) of these
    # energies over short runs: cancellation costs them up to ~5 of the 12
    # digits the energies agree to
    for k in whole["obs"]:
        assert np.allclose(part["obs"][k], whole["obs"][k], rtol=1e-7, atol=1e-12), k
    # the energy the kernel carries (of the last Metropolis sweep) is the
    # system's energy: the overrelaxation sweeps after it conserve it up to
    # FP32 rounding
    assert np.allclose(part["energy"][part["state"]["system_of_slot"]], part["state"]["E"],
                       rtol=1e-5, atol=1e-5 * np.abs(E0).max())
    assert whole["state"]["accepted_exchange"].sum() > 0
