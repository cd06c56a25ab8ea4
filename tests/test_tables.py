"""The product's host table builder (libcsmc csmc_build_tables, the
Lattice constructor of lattice.jl:65-291) against the oracle, bit-exact;
colouring validity; error behaviour of the reference quirks.  No GPU."""
import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from csmc.unitcell import (Honeycomb, Square, addBilinear, addCubic, addOnSite, addQuartic,
                           addZeemanCoupling)


def _cases():
    rng = np.random.default_rng(21)
    out = [
        ("cfg1_square", W.square_heisenberg(), (16, 16)),
        ("cfg2_honeycomb", W.honeycomb_kitaev(1.0, -2.0), (8, 8)),
        ("mctests_kgg", W.kgg_test_cell(), (4, 4)),
        ("cfg3_square_dm", W.square_dm(-1.1, 0.07, [0.1, -0.2, 0.3]), (16, 24)),
        ("cfg3_tri_dm", W.triangular_dm(-0.9, -0.05, [0.3, 0.1, 0.0]), (24, 12)),
        ("cfg4_pyro", W.pyrochlore_mixed(), (3, 3, 3)),
        ("cfg5_tri_j1j2", W.triangular_j1j2(1.0, 0.17), (12, 12)),
        ("tiny_L1", W.square_heisenberg(), (1, 1)),
        ("tiny_L2", W.square_heisenberg(), (2, 2)),
    ]
    uc = Honeycomb()
    addCubic(uc, 2, 1, 2, rng.normal(size=(3, 3, 3)), (1, 0), (0, -1))
    addQuartic(uc, 1, 2, 2, 1, rng.normal(size=(3, 3, 3, 3)), (0, 0), (1, -1), (-1, 1))
    addQuartic(uc, 2, 2, 1, 1, rng.normal(size=(3, 3, 3, 3)), (1, 0), (0, 1), (-1, 0))
    addBilinear(uc, 2, 1, rng.normal(size=(3, 3)), (1, 1))
    addBilinear(uc, 2, 2, rng.normal(size=(3, 3)), (0, 2))
    addOnSite(uc, 1, rng.normal(size=(3, 3)))
    addOnSite(uc, 2, rng.normal(size=(3, 3)))
    addZeemanCoupling(uc, 2, rng.normal(size=3))
    addZeemanCoupling(uc, 1, rng.normal(size=3))
    out.append(("random_terms", uc, (3, 4)))
    return out


@pytest.mark.parametrize("bc", ["periodic", "open"])
def test_tables_match_oracle(bc):
    for name, uc, shape in _cases():
        lat = csmc.Lattice(shape, uc, 1.0, bc=bc, seed=1)
        t = lat.tables()
        o = orc.Tables(uc, shape, bc == "periodic")
        assert np.array_equal(t["pos"], o.pos), name
        assert np.array_equal(t["field"], o.field), name
        assert np.array_equal(t["onsite"], o.onsite), name
        assert np.array_equal(t["bsite"], o.bsite), name
        assert np.array_equal(t["bmat"], o.bmat), name
        assert np.array_equal(t["csite"], o.csite), name
        assert np.array_equal(t["ctens"], o.ctens), name
        assert np.array_equal(t["qsite"], o.qsite), name
        assert np.array_equal(t["qtens"], o.qtens), name
        assert np.array_equal(lat.site_positions, o.pos.T), name


def _conflicts(lat):
    t = lat.tables()
    col = t["colour"]
    N = lat.size
    bad = 0
    for i in range(N):
        members = [int(j) - 1 for j in t["bsite"][i] if j > 0]
        members += [int(j) - 1 for j in t["csite"][i].reshape(-1) if j > 0]
        members += [int(j) - 1 for j in t["qsite"][i].reshape(-1) if j > 0]
        for j in members:
            if j != i and col[j] == col[i]:
                bad += 1
    return bad


def test_colouring_is_independent_sets():
    counts = {}
    for name, uc, shape in _cases():
        lat = csmc.Lattice(shape, uc, 1.0, seed=1)
        assert _conflicts(lat) == 0, name
        counts[name] = lat.info.n_colours
    assert counts["cfg1_square"] == 2
    assert counts["cfg2_honeycomb"] == 2
    assert counts["cfg4_pyro"] == 4  # colour = basis
    assert counts["cfg5_tri_j1j2"] <= 5
    assert counts["cfg3_tri_dm"] == 3  # 3 | L


def test_colour_counts_triangular():
    for L in (16, 24, 32, 48, 64):
        lat = csmc.Lattice((L, L), W.triangular_dm(-1.0, 0.1, [0, 0, 0.1]), 1.0, seed=1)
        assert _conflicts(lat) == 0
        assert lat.info.n_colours == (3 if L % 3 == 0 else 4), L


def test_self_reference_flag():
    lat = csmc.Lattice((4, 4, 4), W.pyrochlore_mixed(), 1.0, seed=1)
    assert lat.info.has_self == 1  # the biquadratic term names the site itself
    lat = csmc.Lattice((8, 8), W.honeycomb_kitaev(1.0, -2.0), 0.5, seed=1)
    assert lat.info.has_self == 0
    lat = csmc.Lattice((1, 1), W.square_heisenberg(), 1.0, seed=1)
    assert lat.info.has_self == 1


def test_reference_error_behaviour():
    uc = Honeycomb()
    addZeemanCoupling(uc, 2, [1.0, 2.0, 3.0])
    with pytest.raises(csmc.BoundsError):
        csmc.Lattice((2, 2), uc, 1.0)
    uc = W.square_heisenberg()
    with pytest.raises(csmc.CsmcError, match="Invalid boundary condition"):
        csmc.Lattice((4, 4), uc, 1.0, bc="twisted")
    uc = Square()  # no terms: the reference never evaluates BC()
    lat = csmc.Lattice((2, 2), uc, 1.0, bc="twisted")
    assert lat.size == 4 and len(uc.basis) == 1  # default basis added to the caller's cell


def test_zero_tensors_dropped():
    uc = Square()
    addBilinear(uc, 1, 1, np.zeros((3, 3)), (1, 0))
    addOnSite(uc, 1, np.zeros((3, 3)))
    addCubic(uc, 1, 1, 1, np.zeros((3, 3, 3)))
    addQuartic(uc, 1, 1, 1, 1, np.zeros((3, 3, 3, 3)))
    addZeemanCoupling(uc, 1, [0.0, 0.0, 0.0])
    assert not uc.bilinear and not uc.onsite and not uc.cubic and not uc.quartic
    assert len(uc.field) == 1


def test_norm_spin_and_params_buffer():
    # test/latticetests.jl:3-7
    lat = csmc.Lattice((2, 2), Square(), 1.0)
    assert np.allclose(np.linalg.norm(lat.spins, axis=0), 1.0)
    d = {"t_thermalization": 5, "bogus": 1}
    with pytest.warns(UserWarning, match="not a valid MC parameter"):
        p = csmc.MCParamsBuffer(d)
    assert d["overrelaxation_rate"] == 10 and p.t_thermalization == 5 and p.report_interval == 0


def test_params_round_trip():
    # UnitCell.params / with_params: the coupling vector in the gradient
    # layout, and a unit cell rebuilt from it equals the original
    from csmc import workloads as W
    uc = W.pyrochlore_mixed()
    uc.field.append((2, np.array([0.1, -0.2, 0.3])))
    theta = uc.params()
    assert theta.size == uc.n_params()
    u2 = uc.with_params(theta)
    for grp in ("field", "onsite", "bilinear", "cubic", "quartic"):
        for a, b in zip(getattr(uc, grp), getattr(u2, grp)):
            for x, y in zip(a, b):
                assert np.array_equal(np.asarray(x), np.asarray(y)), grp
    u3 = uc.with_params(theta * 2.0)
    assert np.array_equal(u3.params(), 2.0 * theta)
    # cubic tensors travel column-major (the csmc_unit_cell layout)
    k = 3 * len(uc.field) + 9 * len(uc.onsite) + 9 * len(uc.bilinear)
    assert np.array_equal(theta[k:k + 27], np.asarray(uc.cubic[0][3]).flatten(order="F"))
