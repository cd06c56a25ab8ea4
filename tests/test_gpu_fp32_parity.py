"""The FP32 headline kernel (register-table storage mode, k_wg_reg) against
the oracle at the headline size: honeycomb Kitaev-Heisenberg J=1, K=-2,
L=64 (8192 spins), S=1/2, 64 replicas, T geometric 0.01..1.

One exchange period of the reference's PT loop (monte_carlo.jl:295-349) is
10 overrelaxation sweeps, one Metropolis sweep with total_energy, and the
exchange.  Its pieces are checked against the oracle's FP64 replay of the
same coloured schedule and Philox stream:
  - overrelaxation (monte_carlo.jl:126-139): from a low-temperature state,
    FP32 spins within 1e-6 S (rms; 5e-6 S at most) of the FP64 replay after
    10 sweeps; from
    random states, where the map is chaotic, every single FP32 reflection
    within 1e-6 S (times the site's conditioning) of the FP64 one on the
    same input;
  - Metropolis (metropolis.jl:65-82): every accept/reject decision equal to
    the FP64 decision, except where the FP64 energy change lies within
    FP32 rounding of the acceptance threshold -T ln u; those cases are
    counted and bounded;
  - total_energy (hamiltonian.jl:70-132) after the Metropolis sweep within
    1e-5 of the FP64 sweep's from the same input;
  - the chained PT form of the same kernel gives the spins of the sweep
    form bit for bit, and its exchange decisions equal the oracle's
    acceptance formula (monte_carlo.jl:308-349) fed the device energies.
"""
import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W

pytestmark = pytest.mark.gpu

TAG_METRO = 0x4D455452
OR_TOL = 1e-6        # one FP32 reflection vs FP64 on the same input, relative to S
OR10_RMS = 1e-6      # ten FP32 sweeps vs FP64 from a low-T state: rms, relative to S
OR10_MAX = 5e-6      # ... and the largest of the 1.6e6 components (rounding of ten
                     # sweeps adds up like a random walk, about 3e-7 per sweep)
E_TOL = 1e-5         # FP32 period energy vs the FP64 trajectory, relative
DE_MARGIN = 1e-5     # FP32 rounding band of the energy change around -T ln u
MAX_MARGINAL = 1e-4  # marginal decisions allowed, as a fraction of the proposals

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox_np(c0, c1, c2, c3, seed):
    """Philox4x32-10 (Salmon et al., SC'11) over numpy arrays of counters;
    pinned to the oracle's known-answer-checked implementation below"""
    c = [np.asarray(x, dtype=np.uint64) & MASK for x in (c0, c1, c2, c3)]
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ np.uint64(k0)) & MASK, p1 & MASK,
             ((p0 >> np.uint64(32)) ^ c[3] ^ np.uint64(k1)) & MASK, p0 & MASK]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c


def u01d(x):
    return (x.astype(np.float64) + 0.5) * 2.3283064365386963e-10


def test_numpy_philox_matches_oracle():
    rng = np.random.default_rng(1)
    for _ in range(20):
        ctr = rng.integers(0, 2 ** 32, size=4, dtype=np.uint64)
        seed = int(rng.integers(0, 2 ** 63))
        got = [int(v) for v in philox_np(*[np.array([x]) for x in ctr], seed)]
        want = list(orc.philox(ctr.astype(np.uint32), [seed & 0xFFFFFFFF, seed >> 32]))
        assert [g for g in got] == [int(w) for w in want]


def _setup(L=64, R=64, seed=5, noise=None):
    """the headline batch; spins random (noise None) or the Klein-dual ground
    state with Gaussian noise of that width per component, renormalised and
    rounded to FP32 (a low-temperature state)"""
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice((L, L), uc, 0.5, seed=seed)
    T = orc.Tables(uc, (L, L))
    temps = W.geometric_temperatures(0.01, 1.0, R)
    b = csmc.SpinBatch([lat] * R, precision=32)
    assert b.storage_mode() == 3, "the headline lattice runs in the register-table mode"
    if noise is None:
        b.randomize(seed)
    else:
        from test_oracle import _klein_state
        g = _klein_state(T, lat.S, uc)
        rng = np.random.default_rng(seed)
        for r in range(R):
            x = g + rng.normal(size=g.shape) * noise * lat.S
            x = lat.S * x / np.linalg.norm(x, axis=1, keepdims=True)
            b.set_spins(r, x.astype(np.float32).astype(np.float64))
    s0 = np.stack([b.get_spins(r) for r in range(R)])  # FP32 values, exactly
    col = lat.colouring()
    classes = [np.nonzero(col == c)[0] + 1 for c in range(int(col.max()) + 1)]
    return lat, T, temps, b, s0, classes


def _proposals(sites1, stream, ctr, seed, S):
    """the FP64 uniform proposal at 1-based sites (random_spin_orientation,
    lattice.jl:306-311) and the uniform of the acceptance test"""
    n = len(sites1)
    r = philox_np(sites1 - 1, np.full(n, stream), np.full(n, ctr), np.full(n, TAG_METRO), seed)
    phi = 2.0 * np.pi * u01d(r[0])
    z = 2.0 * u01d(r[1]) - 1.0
    rr = np.sqrt(1.0 - z * z)
    prop = S * np.stack([rr * np.cos(phi), rr * np.sin(phi), z], axis=1)
    return prop, u01d(r[2])


def test_fp32_overrelaxation_single_updates_from_random_states():
    # From random (infinite-temperature) states the overrelaxation map is
    # chaotic: a 1e-7 perturbation grows about fourfold per sweep and is O(1)
    # after ten, so FP32 and FP64 trajectories separate whatever the kernel
    # does.  What pins the kernel there is each update on its own input: the
    # oracle replays every colour class from the device's state before that
    # class (the sites of a class are independent, monte_carlo.jl:127-137),
    # and each FP32 reflection must equal the FP64 one within its rounding,
    # 1e-6 S times the conditioning kappa = sum_j |J_j s_j| / |H| of the site
    lat, T, temps, b, s0, classes = _setup(seed=7)
    R, S = 8, lat.S
    state = s0.copy()
    worst, frac_over = 0.0, 0.0
    for sweep in range(2):
        b.sweep(csmc.OVERRELAX, n_sweeps=1)
        out = np.stack([b.get_spins(r) for r in range(R)])
        for r in range(R):
            x = state[r].copy()
            for cls in classes:
                H = orc.local_fields(T, x)[cls - 1]
                kappa = 3.0 * S / np.maximum(np.linalg.norm(H, axis=1), 1e-30)
                orc.overrelax(T, x, cls)
                err = np.max(np.abs(out[r][cls - 1] - x[cls - 1]), axis=1) / S
                assert np.all(err <= OR_TOL * np.maximum(1.0, kappa)), (sweep, r, float(err.max()))
                worst = max(worst, float(err.max()))
                frac_over = max(frac_over, float(np.mean(err > OR_TOL)))
                x[cls - 1] = out[r][cls - 1]
            state[r] = out[r]
    assert frac_over < 1e-3, frac_over


def test_fp32_period_vs_oracle_replay():
    # a low-temperature state (the Klein-dual ground state with noise, E/N
    # about -0.345): there the overrelaxation map is regular, and ten FP32
    # sweeps stay within FP32 rounding of the FP64 replay
    lat, T, temps, b, s0, classes = _setup(noise=0.2)
    R, S = len(temps), lat.S
    order = np.concatenate(classes)
    # -- 10 overrelaxation sweeps --------------------------------------------
    b.sweep(csmc.OVERRELAX, n_sweeps=10)
    s1 = np.stack([b.get_spins(r) for r in range(R)])
    worst_or, sq = 0.0, 0.0
    for r in range(R):
        x = s0[r].copy()
        for _ in range(10):
            orc.overrelax(T, x, order)
        worst_or = max(worst_or, float(np.max(np.abs(s1[r] - x))) / S)
        sq += float(np.mean((s1[r] - x) ** 2)) / S ** 2
    assert np.sqrt(sq / R) <= OR10_RMS, np.sqrt(sq / R)
    assert worst_or <= OR10_MAX, worst_or
    # -- one Metropolis sweep --------------------------------------------------
    seed, c0 = 0x5EED0042, 3
    b.set_counters(c0)
    acc_dev = b.sweep(csmc.METROPOLIS, T=temps, n_sweeps=1, seed=seed)
    s2 = np.stack([b.get_spins(r) for r in range(R)])
    E2 = b.total_energy()
    proposals = 0
    marginal = 0
    clean = np.ones(R, dtype=bool)
    for r in range(R):
        taken = 0
        # the oracle decides class by class from the device's own state, so a
        # decision is compared with the FP64 decision at the same input
        state = s1[r].copy()
        for cls in classes:
            _, dec, dE = orc.metropolis_coloured(T, state, cls, S, temps[r], 0, 60.0, seed, r, c0)
            prop, u = _proposals(cls, r, c0, seed, S)
            dev = s2[r][cls - 1]
            old = s1[r][cls - 1]
            took = np.linalg.norm(dev - prop, axis=1) < np.linalg.norm(dev - old, axis=1)
            differ = np.nonzero(took != dec.astype(bool))[0]
            if len(differ):
                clean[r] = False
                crit = -temps[r] * np.log(u[differ])
                band = DE_MARGIN * np.maximum(1.0, np.abs(dE[differ]))
                assert np.all(np.abs(dE[differ] - crit) <= band), (r, differ, dE[differ], crit)
                marginal += len(differ)
            proposals += len(cls)
            taken += int(took.sum())
            # the next class starts from the device's result for this class
            state[cls - 1] = dev
        assert int(acc_dev[r]) == taken, (r, acc_dev[r], taken)
    assert marginal <= MAX_MARGINAL * proposals, (marginal, proposals)
    # -- the period's energy: the FP64 Metropolis sweep from the same input ----
    for r in range(R):
        if clean[r]:
            x = s1[r].copy()
            orc.metropolis_coloured(T, x, order, S, temps[r], 0, 60.0, seed, r, c0)
            Eo = orc.total_energy(T, x)
            assert abs(E2[r] - Eo) <= E_TOL * abs(Eo), (r, E2[r], Eo)
            assert np.max(np.abs(s2[r] - x)) <= 1e-5 * S, r
    assert clean.sum() >= R - 4, f"{R - clean.sum()} replicas with marginal decisions"


def test_fp32_pt_form_equals_sweep_form_and_exchanges():
    # parallel tempering sweeps 0..20: sweeps 0, 10 and 20 run overrelaxation,
    # Metropolis + total_energy and the exchange (swap_rate = or_rate = 10);
    # the chained PT kernel must equal the sweep form bit for bit, and each
    # exchange must be the oracle's formula on the device's energies
    lat, T, temps, b, s0, classes = _setup(seed=6)
    R = len(temps)
    seed = 0xA11CE
    p = dict(t_thermalization=1000, t_measurement=0, probe_rate=10, swap_rate=10, overrelaxation_rate=10,
             report_interval=10 ** 6, checkpoint_rate=10 ** 6)
    pt = csmc.SpinBatch([lat] * R, precision=32)
    for r in range(R):
        pt.set_spins(r, s0[r])
    pt.set_counters(0)
    pt.pt_init(temps, p, csmc.METROPOLIS, seed=seed)
    b.set_counters(0)
    perm = np.arange(R, dtype=np.intc)      # slot -> system, oracle side
    acc_x = np.zeros(R, dtype=np.int_)
    for xc, until in enumerate((1, 11, 21)):
        stopped, _ = pt.pt_advance(until)
        assert stopped == until
        # the sweep form of the same sweeps: each system at its slot's T
        T_sys = np.empty(R)
        T_sys[perm] = temps
        n_or = 1 if xc == 0 else 10
        b.sweep(csmc.OVERRELAX, n_sweeps=n_or)
        b.sweep(csmc.METROPOLIS, T=T_sys, n_sweeps=1, seed=seed)
        E_sys = b.total_energy()
        for r in range(R):
            assert np.array_equal(pt.get_spins(r), b.get_spins(r)), (xc, r)
        Eslot = np.ascontiguousarray(E_sys[perm])
        orc.pt_exchange(temps, Eslot, perm, seed, xc, acc_x)
        st = pt.pt_state()
        assert np.array_equal(st["system_of_slot"], perm), xc
        assert np.array_equal(st["accepted_exchange"], acc_x), xc
        assert np.allclose(st["E"], E_sys[perm], rtol=0, atol=1e-9 * np.max(np.abs(E_sys)))
    assert acc_x.sum() > 0
