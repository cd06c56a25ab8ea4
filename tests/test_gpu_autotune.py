"""csmc_autotune: every feasible sweep form is timed on the batch's own
state, the state is restored afterwards, and the sweeps that follow (in the
chosen form) make the same updates as an untuned twin batch: FP64 spins to
1e-12 and identical Metropolis decisions."""
import numpy as np
import pytest

import csmc
from csmc import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [csmc.OVERRELAX, csmc.METROPOLIS], ids=["overrelax", "metropolis"])
def test_autotune_restores_state_and_keeps_results(kind):
    lats = [csmc.Lattice((16, 16), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=3 + k) for k in range(4)]
    T = np.array([0.2, 0.4, 0.8, 1.6])
    tuned = csmc.SpinBatch(lats, precision=64)
    twin = csmc.SpinBatch(lats, precision=64)
    for b in (tuned, twin):
        b.randomize(9)
    before = [tuned.get_spins(r) for r in range(4)]
    ctr = tuned.counters()
    times, choice = tuned.autotune(kind, T=T, n_sweeps=2)
    print(f"[autotune] {times} -> {choice}", flush=True)
    assert choice in times and len(times) >= 2
    assert all(t > 0 for t in times.values())
    for r in range(4):
        assert np.array_equal(tuned.get_spins(r), before[r])
    assert np.array_equal(tuned.counters(), ctr)
    for _ in range(2):
        a1 = tuned.sweep(kind, T=T, sigma=0.9, n_sweeps=3, seed=17)
        a2 = twin.sweep(kind, T=T, sigma=0.9, n_sweeps=3, seed=17)
        for r in range(4):
            x, y = tuned.get_spins(r), twin.get_spins(r)
            assert np.max(np.abs(x - y)) <= 1e-12, r
        assert np.array_equal(a1, a2)
        assert np.array_equal(tuned.counters(), twin.counters())
    tuned.close()
    twin.close()
