"""The oracle's replay of the reference's PT loop over coloured sweeps
(monte_carlo.jl:295-349), shared by the GPU parity tests.  Test helper: it
drives the oracle (the checker), never the product."""
import numpy as np

import oracle.oracle as orc


def colour_order(lat):
    """1-based sites, colour class by colour class (the device's sweep order)"""
    col = lat.colouring()
    order = []
    for c in range(int(col.max()) + 1):
        order.extend((np.nonzero(col == c)[0] + 1).tolist())
    return np.array(order)


def pt_replay(T, spins, order, temps, S, p, seed, alg_kind, snapshots=(), n_sweeps=None):
    """Returns (spins per system, slot->system, E per system, acc_local per
    slot, acc_exch per slot, {sweep: (spins per slot)}).  A snapshot at sweep
    s is taken after that sweep's exchange, where the reference writes its
    checkpoints (monte_carlo.jl:353-365)."""
    R = len(temps)
    perm = np.arange(R, dtype=np.intc)
    E = np.array([orc.total_energy(T, spins[r]) for r in range(R)])
    acc_l = np.zeros(R)
    acc_x = np.zeros(R, dtype=np.int_)
    mcount = np.zeros(R, dtype=np.int64)
    dosweep = p["overrelaxation_rate"] or 1
    total = p["t_thermalization"] + p["t_measurement"] if n_sweeps is None else n_sweeps
    snaps = {}
    for sw in range(total):
        for r in range(R):  # r = system
            slot = int(np.nonzero(perm == r)[0][0])
            if p["overrelaxation_rate"]:
                orc.overrelax(T, spins[r], order)
            if sw % dosweep == 0:
                a, _, _ = orc.metropolis_coloured(T, spins[r], order, S, temps[slot], alg_kind, 60.0, seed, r,
                                                  int(mcount[r]))
                mcount[r] += 1
                acc_l[slot] += a
                E[r] = orc.total_energy(T, spins[r])
        if R > 1 and sw % dosweep == 0 and sw % p["swap_rate"] == 0:
            Eslot = np.ascontiguousarray(E[perm])
            orc.pt_exchange(temps, Eslot, perm, seed, sw // p["swap_rate"], acc_x)
        if sw in snapshots:
            snaps[sw] = np.stack([spins[perm[r]].copy() for r in range(R)])
    return spins, perm, E, acc_l, acc_x, snaps
