"""HIP path vs the oracle, FP64: local field, site energy and total energy to
1e-12 relative; coloured overrelaxation / deterministic passes to 1e-12;
Metropolis accept/reject decisions identical to the oracle's replay of the
same coloured schedule and Philox stream; parallel-tempering exchange
replayed step by step."""
import math

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from csmc.unitcell import Honeycomb, addBilinear, addCubic, addOnSite, addQuartic, addZeemanCoupling
from conftest import random_spins
from pt_replay import pt_replay

pytestmark = pytest.mark.gpu

RTOL = 1e-12  # FP64 parity bound stated for this path


def _random_cell():
    rng = np.random.default_rng(21)
    uc = Honeycomb()
    addCubic(uc, 2, 1, 2, rng.normal(size=(3, 3, 3)), (1, 0), (0, -1))
    addQuartic(uc, 1, 2, 2, 1, rng.normal(size=(3, 3, 3, 3)), (0, 0), (1, -1), (-1, 1))
    addBilinear(uc, 2, 1, rng.normal(size=(3, 3)), (1, 1))
    addBilinear(uc, 2, 2, rng.normal(size=(3, 3)), (0, 2))
    addOnSite(uc, 1, rng.normal(size=(3, 3)))
    addOnSite(uc, 2, rng.normal(size=(3, 3)))
    addZeemanCoupling(uc, 2, rng.normal(size=3))
    addZeemanCoupling(uc, 1, rng.normal(size=3))
    return uc


CASES = [
    ("cfg1_square", lambda: W.square_heisenberg(), (16, 16), 1.0, "periodic"),
    ("cfg2_honeycomb", lambda: W.honeycomb_kitaev(1.0, -2.0), (64, 64), 0.5, "periodic"),
    ("mctests_kgg", W.kgg_test_cell, (4, 4), 1.0, "periodic"),
    ("cfg3_square_dm", lambda: W.square_dm(-1.1, 0.07, [0.1, -0.2, 0.3]), (16, 24), 1.0, "periodic"),
    ("cfg3_tri_dm", lambda: W.triangular_dm(-0.9, -0.05, [0.3, 0.1, 0.0]), (24, 16), 1.0, "periodic"),
    ("cfg4_pyro", lambda: W.pyrochlore_mixed(), (4, 4, 4), 1.0, "periodic"),
    ("cfg5_tri_j1j2", lambda: W.triangular_j1j2(1.0, 0.13), (24, 24), 1.0, "periodic"),
    ("tensor_patterns", W.tensor_pattern_cell, (4, 4), 1.0, "periodic"),
    ("random_terms", _random_cell, (3, 4), 1.0, "periodic"),
    ("open_kgg", W.kgg_test_cell, (5, 3), 1.0, "open"),
    ("tiny_L1", lambda: W.square_heisenberg(), (1, 1), 1.0, "periodic"),
]
IDS = [c[0] for c in CASES]


def _setup(case, seed=3):
    name, mk, shape, S, bc = case
    uc = mk()
    lat = csmc.Lattice(shape, uc, S, bc=bc, seed=seed)
    T = orc.Tables(uc, shape, bc == "periodic")
    return lat, T


def _close(a, b, rtol=RTOL):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    return np.max(np.abs(a - b)) <= rtol * scale if b.size else True


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("prec", [64, 32])
def test_evaluation_parity(case, prec):
    lat, T = _setup(case)
    b = csmc.SpinBatch([lat], precision=prec)
    s = lat._spins.copy()
    if prec == 32:
        s = s.astype(np.float32).astype(np.float64)  # the state the FP32 batch holds
    b.set_spins(0, s)
    H = b.local_field(0)
    e = b.site_energy(0)
    E, parts = b.total_energy(parts=True)
    Ho = orc.local_fields(T, s)
    eo = orc.site_energies(T, s)
    Eo, po = orc.total_energy(T, s, parts=True)
    assert _close(H, Ho), np.max(np.abs(H - Ho))
    assert _close(e, eo), np.max(np.abs(e - eo))
    # total energy: relative to the sum of the magnitudes of its parts
    assert abs(E[0] - Eo) <= RTOL * max(1.0, float(np.sum(np.abs(po)))), (E[0], Eo)
    assert _close(parts[0], po)
    assert np.array_equal(b.get_spins(0), s)


def _colour_order(lat):
    col = lat.colouring()
    order = []
    for c in range(int(col.max()) + 1):
        order.extend((np.nonzero(col == c)[0] + 1).tolist())
    return np.array(order)


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("plain", [False, True])
def test_overrelaxation_and_deterministic_parity(case, plain):
    lat, T = _setup(case)
    b = csmc.SpinBatch([lat], precision=64)
    s = lat._spins.copy()
    b.set_spins(0, s)
    order = _colour_order(lat)
    b.sweep(csmc.OVERRELAX, n_sweeps=3, plain=plain)
    for _ in range(3):
        orc.overrelax(T, s, order)
    assert _close(b.get_spins(0), s, 1e-11)
    b.sweep(csmc.DETERMINISTIC, n_sweeps=2, plain=plain)
    for _ in range(2):
        orc.deterministic(T, s, order, lat.S)
    assert _close(b.get_spins(0), s, 1e-11)


@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("kind", [csmc.METROPOLIS, csmc.CONE, csmc.CONE_ADAPTIVE])
@pytest.mark.parametrize("plain", [False, True])
def test_metropolis_decisions_identical(case, kind, plain):
    lat, T = _setup(case)
    b = csmc.SpinBatch([lat], precision=64)
    s = lat._spins.copy()
    b.set_spins(0, s)
    c0 = 17
    b.set_counters(c0)
    order = _colour_order(lat)
    seed = 0x1234ABCD5678
    temp = 0.7
    sigma = np.array([0.4])
    sig_o = 0.4
    acc_dev, acc_orc = 0.0, 0
    for sw in range(4):
        acc_dev += float(b.sweep(kind, T=temp, sigma=sigma, n_sweeps=1, seed=seed, plain=plain)[0])
        a, dec, dE = orc.metropolis_coloured(T, s, order, lat.S, temp, 0 if kind == csmc.METROPOLIS else 1,
                                             sig_o, seed, 0, c0 + sw)
        acc_orc += a
        if kind == csmc.CONE_ADAPTIVE:
            sig_o = orc.adapt_sigma(sig_o, a, T.N)
    assert acc_dev == acc_orc
    assert _close(b.get_spins(0), s, 1e-10)
    if kind == csmc.CONE_ADAPTIVE:
        assert abs(sigma[0] - sig_o) <= 1e-12 * sig_o
    assert int(b.counters()[0]) == c0 + 4
    assert 0 < acc_orc < 4 * T.N or T.N == 1


def _pt_replay(T, spins, order, temps, S, p, seed, alg_kind):
    return pt_replay(T, spins, order, temps, S, p, seed, alg_kind)[:5]


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[5]], ids=[IDS[0], IDS[2], IDS[5]])
def test_parallel_tempering_replay(case):
    lat, T = _setup(case)
    R = 6
    rng = np.random.default_rng(8)
    spins = np.stack([random_spins(lat.size, lat.S, rng) for _ in range(R)])
    temps = W.geometric_temperatures(0.3, 3.0, R)
    # 16 sweeps: overrelaxation is chaotic at high T, so last-bit differences
    # between device and host sin/cos grow ~2x per sweep; 16 sweeps keep the
    # FP64 trajectories within 1e-10 while every schedule branch is visited
    p = dict(t_thermalization=7, t_measurement=9, probe_rate=3, swap_rate=4, overrelaxation_rate=2,
             report_interval=7, checkpoint_rate=5)
    seed = 99
    b = csmc.SpinBatch([lat] * R, precision=64)
    for r in range(R):
        b.set_spins(r, spins[r])
    b.set_counters(0)
    b.pt_init(temps, p, csmc.METROPOLIS, seed=seed)
    events = []
    while True:
        stopped, ev = b.pt_advance(-1)
        events.append((stopped, ev))
        if ev == 3:
            break
    st = b.pt_state()
    order = _colour_order(lat)
    so, perm, E, acc_l, acc_x = _pt_replay(T, spins.copy(), order, temps, lat.S, p, seed, 0)
    assert np.array_equal(st["system_of_slot"], perm)
    assert np.array_equal(st["accepted_local"], acc_l)
    assert np.array_equal(st["accepted_exchange"], acc_x)
    assert _close(st["E"], E[perm], 1e-10)
    # per-spin bound: the pyrochlore model's overrelaxation (onsite, cubic,
    # quartic terms) is not energy conserving and amplifies last-bit
    # differences fastest; the discrete schedule above is checked exactly
    for r in range(R):
        assert _close(b.get_spins(r), so[r], 1e-7)
    # host events: checkpoints at sweeps >= 7 divisible by 5, reports at 7, 14
    assert (10, 1) in events and (15, 1) in events and (7, 2) in events and (14, 2) in events
    assert events[-1] == (16, 3)


def test_fp32_random_spins_uniform():
    # the FP32 proposal uses the hardware sin/cos in turns: the directions
    # must be uniform on the sphere (mean 0, <s_a^2> = S^2/3)
    lat = csmc.Lattice((64, 64), W.square_heisenberg(), 1.0, seed=0)
    b = csmc.SpinBatch([lat] * 4, precision=32)
    b.randomize(2024)
    s = np.concatenate([b.get_spins(i) for i in range(4)])
    n = len(s)
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=2e-6)
    assert np.all(np.abs(s.mean(axis=0)) < 5 / math.sqrt(3 * n))
    assert np.all(np.abs((s ** 2).mean(axis=0) - 1 / 3) < 5 * math.sqrt(4 / 45 / n))
