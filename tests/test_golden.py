"""The committed golden vectors (tests/golden/) against the oracle and the
library's host-side lattice builder; no device needed.  The vectors were
written by tests/golden/make_golden.py from the oracle, so this pins the
checker against silent change; the reference's own known answers
(tests/golden/reference_kats.json) pin the library's host tables and the
oracle's energies."""
import json
import os

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc.unitcell import Square, addBilinear, addZeemanCoupling
from golden_cases import CASES, colour_order, make_cell

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(HERE, "golden.npz"))
SEED = int(G["params"][0]) | (int(G["params"][1]) << 32)
CTR0 = int(G["params"][2])
TEMP, SIGMA = (float(v) for v in G["params_f"])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_reproduces_golden(case):
    name, mk, shape, S, bc = case
    p = name + "/"
    uc = make_cell(mk)
    T = orc.Tables(uc, shape, bc == "periodic")
    s = G[p + "spins"].copy()
    order = G[p + "order"]
    assert np.array_equal(T.bsite, G[p + "bsite"])
    assert np.array_equal(orc.local_fields(T, s), G[p + "local_field"])
    assert np.array_equal(orc.site_energies(T, s), G[p + "site_energy"])
    E, parts = orc.total_energy(T, s, parts=True)
    assert E == G[p + "total_energy"][0] and np.array_equal(parts, G[p + "energy_parts"])
    so = s.copy()
    for _ in range(2):
        orc.overrelax(T, so, order)
    assert np.array_equal(so, G[p + "overrelax2"])
    for kind, tag in ((0, "metropolis"), (1, "cone")):
        sm = s.copy()
        for sw in range(2):
            a, dec, _ = orc.metropolis_coloured(T, sm, order, S, TEMP, kind, SIGMA, SEED, 0, CTR0 + sw)
            assert a == G[p + tag + "_accepted"][sw]
            assert np.array_equal(dec, G[p + tag + "_decisions"][sw])
        assert np.array_equal(sm, G[p + tag + "_spins"])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_host_tables_match_golden(case):
    # the library's host builder (csmc_build_tables) gives the reference's
    # bilinear table and the colour order the sweep vectors were made with
    name, mk, shape, S, bc = case
    lat = csmc.Lattice(shape, mk(), S, bc=bc, seed=11)
    assert np.array_equal(lat.tables()["bsite"], G[name + "/bsite"])
    assert np.array_equal(colour_order(lat), G[name + "/order"])
    assert np.array_equal(lat._spins, G[name + "/spins"])


def test_reference_known_answers():
    with open(os.path.join(HERE, "reference_kats.json")) as f:
        kats = json.load(f)
    # test/latticetests.jl:9-18
    uc = Square()
    addZeemanCoupling(uc, 1, [1.0, 0.0, 0.0])
    uc.basis.append(np.zeros(2))
    T = orc.Tables(uc, (1, 1))
    s = np.array([[1.0, 0.0, 0.0]])
    k = kats["zeeman_single_site"]
    assert orc.total_energy(T, s) == k["E"]
    assert np.array_equal(orc.local_field(T, s, 1), k["H"])
    # test/latticetests.jl:20-30
    uc = Square()
    for d in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        addBilinear(uc, 1, 1, -np.eye(3), d)
    lat = csmc.Lattice((2, 2), uc, 1.0, seed=0)
    T = orc.Tables(uc, (2, 2))
    s = np.tile([1.0, 0.0, 0.0], (4, 1))
    assert orc.total_energy(T, s) / 4 == kats["square_fm_L2"]["E_per_site"]
    # test/latticetests.jl:6
    assert np.allclose(np.linalg.norm(lat._spins, axis=1), kats["spin_norm"]["norm"], atol=1e-9)
