"""The HIP path against the committed golden vectors (tests/golden/), FP64,
through the C ABI: evaluation to 1e-12 relative, overrelaxation and
deterministic sweeps to 1e-11, Metropolis and cone-move sweeps with the
same accepted counts and final spins as the golden replay of the coloured
schedule and Philox stream."""
import os

import numpy as np
import pytest

import csmc
from golden_cases import CASES, colour_order

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(HERE, "golden.npz"))
SEED = int(G["params"][0]) | (int(G["params"][1]) << 32)
CTR0 = int(G["params"][2])
TEMP, SIGMA = (float(v) for v in G["params_f"])


def _close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= rtol * max(1.0, float(np.max(np.abs(b))))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_hip_path_matches_golden(case):
    name, mk, shape, S, bc = case
    p = name + "/"
    lat = csmc.Lattice(shape, mk(), S, bc=bc, seed=11)
    assert np.array_equal(colour_order(lat), G[p + "order"])
    s = G[p + "spins"]
    b = csmc.SpinBatch([lat], precision=64)
    b.set_spins(0, s)
    assert _close(b.local_field(0), G[p + "local_field"], 1e-12)
    assert _close(b.site_energy(0), G[p + "site_energy"], 1e-12)
    E, parts = b.total_energy(parts=True)
    assert abs(E[0] - G[p + "total_energy"][0]) <= 1e-12 * max(1.0, float(np.abs(G[p + "energy_parts"]).sum()))
    b.sweep(csmc.OVERRELAX, n_sweeps=2)
    assert _close(b.get_spins(0), G[p + "overrelax2"], 1e-11)
    b.set_spins(0, s)
    b.sweep(csmc.DETERMINISTIC, n_sweeps=1)
    assert _close(b.get_spins(0), G[p + "deterministic1"], 1e-11)
    for kind, tag in ((csmc.METROPOLIS, "metropolis"), (csmc.CONE, "cone")):
        b.set_spins(0, s)
        b.set_counters(CTR0)
        for sw in range(2):
            acc = b.sweep(kind, T=TEMP, sigma=np.array([SIGMA]), n_sweeps=1, seed=SEED)
            assert int(acc[0]) == int(G[p + tag + "_accepted"][sw]), (tag, sw)
        assert _close(b.get_spins(0), G[p + tag + "_spins"], 1e-10)
    b.close()
