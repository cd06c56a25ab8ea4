"""metropolis_constraint_adaptive! (metropolis.jl:155-179) on the device for a
site-sum constraint (SiteSumConstraint) against the oracle's literal
restatement (orc_metropolis_constraint_coloured: the constraint evaluated
over the whole lattice with the proposed spin in place, per proposal),
replaying the same coloured schedule and Philox stream: identical decisions
and accepted counts, spins to 1e-10 (FP64), the multiplier lambda and the
cone width sigma (width law :176-177) per sweep."""
import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from pt_replay import colour_order

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "honeycomb_mz":
        uc = W.honeycomb_kitaev(1.0, -2.0, h=[0.05, 0.0, 0.1])
        a = [[0, 0, 1], [0, 0, 1]]
        Q = np.zeros((2, 3, 3))
        return uc, (6, 6), 0.5, a, Q, 4.0, 0.02
    if name == "staggered_quadratic":
        uc = W.kgg_test_cell()
        rng = np.random.default_rng(3)
        q = 0.1 * rng.normal(size=(3, 3))
        return uc, (4, 4), 1.0, [[0.3, 0, 1], [0, 0.2, -1]], np.stack([q + q.T, q @ q.T]), None, 0.1
    uc = W.pyrochlore_mixed()  # cubic, quartic and onsite terms; self references
    return uc, (2, 2, 2), 1.0, [[0, 0, 1]] * 4, np.zeros((4, 3, 3)), None, 0.05


@pytest.mark.parametrize("name", ["honeycomb_mz", "staggered_quadratic", "pyrochlore"])
def test_constraint_sweeps_replay(name):
    uc, shape, S, a, Q, c0, w = _case(name)
    lat = csmc.Lattice(shape, uc, S, seed=4)
    T = orc.Tables(uc, shape)
    R = 3
    temps = np.array([0.2, 0.6, 1.5])
    b = csmc.SpinBatch([lat] * R, precision=64)
    b.randomize(9)
    start = [b.get_spins(r) for r in range(R)]
    if c0 is None:
        # a target near the start's value, so that proposals are accepted
        probe = csmc.Lattice(shape, uc, S, seed=4)
        probe._spins[:] = start[0]
        c0 = csmc.SiteSumConstraint(a, Q, 0.0)(probe) + 0.5
    b.set_counters(5)
    b.set_constraint(a, Q, c0, w, sigma0=2.0)
    lam0 = np.array([0.0, 0.3, -0.2])
    b.constraint_lambda(lam0)
    sig = np.array([0.5, 0.8, 1.1])
    seed = 0xC0FFEE
    order = colour_order(lat)
    ncell = lat.size // len(uc.basis)
    ref = [s.copy() for s in start]
    lam_o = lam0.copy()
    sig_o = sig.copy()
    for sw in range(4):
        acc = b.sweep(csmc.CONSTRAINT_ADAPTIVE, T=temps, sigma=sig, n_sweeps=1, seed=seed)
        lam = b.constraint_lambda()
        for r in range(R):
            n, dec, sig_o[r], lam_o[r] = orc.metropolis_constraint_coloured(
                T, ref[r], order, S, temps[r], sig_o[r], 2.0, seed, r, 5 + sw, ncell,
                np.asarray(a, float), np.asarray(Q, float).reshape(-1, 9), c0, w, lam_o[r])
            assert int(acc[r]) == n, (sw, r, acc[r], n)
            assert np.max(np.abs(b.get_spins(r) - ref[r])) <= 1e-10 * S, (sw, r)
            assert abs(lam[r] - lam_o[r]) <= 1e-10 * max(1.0, abs(lam_o[r])), (sw, r, lam[r], lam_o[r])
            assert abs(sig[r] - sig_o[r]) <= 1e-12 * sig_o[r], (sw, r, sig[r], sig_o[r])
    assert np.all(b.counters() == 9)
    assert np.all(lam != lam0), "every system accepted proposals and moved its multiplier"


def test_constraint_fp32_state_statistically_near_fp64():
    # the FP32 state runs the same kernels with FP32 proposals and energy
    # changes (decisions in FP64); after a few sweeps its multiplier and
    # constraint value stay close to the FP64 run's
    uc, shape, S, a, Q, c0, w = _case("honeycomb_mz")
    lat = csmc.Lattice(shape, uc, S, seed=4)
    out = {}
    for prec in (32, 64):
        b = csmc.SpinBatch([lat] * 2, precision=prec)
        b.randomize(9)
        b.set_constraint(a, Q, c0, w, sigma0=2.0)
        b.sweep(csmc.CONSTRAINT_ADAPTIVE, T=[0.5, 0.5], sigma=np.array([0.5, 0.5]), n_sweeps=1, seed=1)
        out[prec] = (b.constraint_lambda(), b.get_spins(0))
    assert np.allclose(np.linalg.norm(out[32][1], axis=1), S, atol=1e-6)
    assert np.allclose(out[32][0], out[64][0], rtol=1e-3, atol=1e-3)


def test_annealing_with_the_constraint_plugin():
    # simulated_annealing! (monte_carlo.jl:166-187) with
    # MetropolisConstraintAdaptive: overrelaxation every step, the
    # constraint sweep every or_rate-th step, sigma reset per temperature;
    # replayed by the oracle (colour-ordered overrelaxation + the
    # constraint chain), lambda carried across temperatures
    uc, shape, S, a, Q, c0, w = _case("honeycomb_mz")
    lat = csmc.Lattice(shape, uc, S, seed=7)
    T = orc.Tables(uc, shape)
    order = colour_order(lat)
    ncell = lat.size // 2
    # two temperatures of three sweep steps: overrelaxation from a random
    # state is chaotic (last-bit differences of device and host sin/cos grow
    # with every sweep), so the replay is kept to eight sweeps
    mc = csmc.MonteCarlo(0.3, lat, {"t_thermalization": 4, "overrelaxation_rate": 3}, seed=12, precision=64,
                         constraint=csmc.SiteSumConstraint(a, Q, c0), weight=w, sigma0=1.5)
    s = mc.lattice._spins.copy()
    csmc.simulated_annealing(mc, lambda x: 0.5 ** x, 1.0, alg=csmc.MetropolisConstraintAdaptive())
    temps = [1.0, 0.5]
    lam, ctr = 0.0, 0
    for Tk in temps:
        sig = 1.5
        for t in range(1, 4):
            orc.overrelax(T, s, order)
            if t % 3 == 0:
                _, _, sig, lam = orc.metropolis_constraint_coloured(T, s, order, S, Tk, sig, 1.5, 12, 0, ctr, ncell,
                                                                    np.asarray(a, float), np.zeros((2, 9)), c0, w,
                                                                    lam)
                ctr += 1
    assert ctr == 2
    assert np.max(np.abs(mc.lattice._spins - s)) <= 1e-10
    assert abs(mc.lambda_ - lam) <= 1e-10 * max(1.0, abs(lam)) and lam != 0.0
    # the host-side closure form of the constraint agrees with the device's
    assert abs(mc.constraint(mc.lattice) - (np.sum(s[:, 2]) - c0)) < 1e-9


def test_constraint_plugin_refusals():
    lat = csmc.Lattice((4, 4), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=1)
    b = csmc.SpinBatch([lat], precision=64)
    with pytest.raises(csmc.CsmcError, match="no constraint"):
        b.sweep(csmc.CONSTRAINT_ADAPTIVE, T=[0.5], n_sweeps=1)
    mc = csmc.MonteCarlo(0.1, lat, {"t_thermalization": 5}, constraint=lambda L: 0.0, weight=0.1)
    with pytest.raises(TypeError, match="SiteSumConstraint"):
        csmc.simulated_annealing(mc, lambda x: 0.5 ** x, 1.0, alg=csmc.MetropolisConstraintAdaptive())
    with pytest.raises(TypeError, match="no device form"):
        csmc.simulated_annealing(mc, lambda x: 0.5 ** x, 1.0, alg=lambda m, t: 0.0)
    with pytest.raises(NotImplementedError, match="e_diff"):
        csmc.MetropolisConstraint()
