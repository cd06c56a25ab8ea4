"""Per-site queries get_local_field(lattice, point) and energy(lattice,
point) (hamiltonian.jl:3-67,139-196) through csmc_site_eval: only the site
and its partners are written to the device before the evaluation, so a call
does not depend on the lattice size.  Each call must equal the whole-lattice
evaluation bit for bit and the oracle to 1e-12, also after the host spins
changed between calls."""
import time

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from csmc.lattice import energy, get_local_field, local_fields, site_energies

pytestmark = pytest.mark.gpu

CASES = [
    ("pyrochlore_mixed_4", lambda: W.pyrochlore_mixed(), (4, 4, 4), 1.0),
    ("honeycomb_field_8", lambda: W.honeycomb_kitaev(1.0, -2.0, h=(0.1, 0.2, 0.3)), (8, 8), 0.5),
    ("tensor_pattern_6", lambda: W.tensor_pattern_cell(), (6, 6), 1.0),
]


@pytest.mark.parametrize("name,make,shape,S", CASES, ids=[c[0] for c in CASES])
def test_site_queries_equal_whole_lattice_and_oracle(name, make, shape, S):
    uc = make()
    lat = csmc.Lattice(shape, uc, S, seed=3)
    T = orc.Tables(uc, shape)
    rng = np.random.default_rng(4)
    sites = rng.choice(lat.size, size=min(lat.size, 40), replace=False)
    for rnd in range(3):
        if rnd:
            # change some spins on the host between the calls
            for s in rng.choice(lat.size, size=max(1, lat.size // 8), replace=False):
                v = rng.normal(size=3)
                lat._spins[s] = S * v / np.linalg.norm(v)
        H_all = local_fields(lat)
        e_all = site_energies(lat)
        H_orc = orc.local_fields(T, lat._spins)
        for s in sites:
            H = np.array(get_local_field(lat, int(s) + 1))
            assert np.array_equal(H, H_all[s]), (name, rnd, s)
            assert np.max(np.abs(H - H_orc[s])) <= 1e-12 * max(1.0, np.max(np.abs(H_orc[s])))
            assert energy(lat, int(s) + 1) == e_all[s]
        # a site queried right after its neighbour changed
        s = int(sites[0])
        j = int(lat._site_and_partners(s)[-1])
        lat._spins[j] = -lat._spins[j]
        assert np.array_equal(np.array(get_local_field(lat, s + 1)), local_fields(lat)[s])


def test_site_query_bounds():
    lat = csmc.Lattice((4, 4), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=1)
    with pytest.raises(IndexError):
        get_local_field(lat, lat.size + 1)
    with pytest.raises(IndexError):
        energy(lat, 0)


def test_site_query_cost_does_not_grow_with_the_lattice():
    """a query on a 64x larger lattice costs about the same (the whole-lattice
    evaluation would copy and evaluate 64x the sites)"""
    uc = W.honeycomb_kitaev(1.0, -2.0)

    def per_call(L):
        lat = csmc.Lattice((L, L), uc, 0.5, seed=1)
        get_local_field(lat, 1)
        t0 = time.perf_counter()
        for k in range(200):
            get_local_field(lat, 1 + (k * 37) % lat.size)
        return (time.perf_counter() - t0) / 200

    small, large = per_call(16), per_call(128)
    assert large < 4 * small, (small, large)
