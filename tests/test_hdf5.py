"""HDF5 layout (hdf5.jl) round trips, mirroring test/h5tests.jl:5-46 and
extending it to non-empty term groups.  No GPU: the writer works on host
data."""
import os

import numpy as np
import pytest

import csmc
from csmc import hdf5io as H
from csmc import workloads as W


def _same_uc(a, b):
    assert all(np.array_equal(x, y) for x, y in zip(a.lattice_vectors, b.lattice_vectors))
    assert all(np.array_equal(x, y) for x, y in zip(a.basis, b.basis))

    def norm(terms):
        out = []
        for t in terms:
            out.append(tuple(np.asarray(v).tolist() if isinstance(v, np.ndarray) else v for v in t))
        return sorted(out, key=repr)

    for grp in ("field", "onsite", "bilinear", "cubic", "quartic"):
        assert norm(getattr(a, grp)) == norm(getattr(b, grp)), grp


def _key_sorted(uc):
    """the unit cell with every term group in the order read_unit_cell
    returns it: HDF5 lists group members by name (as HDF5.jl's keys() does),
    so a reloaded cell has its terms sorted by their on-disk key"""
    j = H._jl_tuple
    uc.field.sort(key=lambda t: str(t[0]))
    uc.onsite.sort(key=lambda t: str(t[0]))
    uc.bilinear.sort(key=lambda t: f"({t[0]},{t[1]}),{j(t[3])}")
    uc.cubic.sort(key=lambda t: f"({t[0]},{t[1]},{t[2]}),{j(t[4])},{j(t[5])}")
    uc.quartic.sort(key=lambda t: f"({t[0]},{t[1]},{t[2]},{t[3]}),{j(t[5])},{j(t[6])},{j(t[7])}")
    return uc


@pytest.mark.parametrize("name", ["square_bare", "pyrochlore_mixed", "kgg"])
def test_metadata_round_trip(tmp_path, name):
    # test/h5tests.jl: dump_metadata! then read_unit_cell / read_lattice
    uc = {"square_bare": csmc.Square, "pyrochlore_mixed": W.pyrochlore_mixed,
          "kgg": W.kgg_test_cell}[name]()
    uc = _key_sorted(uc)
    shape = (2, 2, 2) if uc.D == 3 else (2, 2)
    lat = csmc.Lattice(shape, uc, 1.0, seed=1)
    params = {"t_thermalization": 10 ** 4, "overrelaxation_rate": 10, "t_deterministic": 10 ** 7}
    mc = csmc.MonteCarlo(1.0, lat, params)
    path = str(tmp_path / ".test.h5")
    with H.h5open(path, "w") as f:
        H.dump_metadata(f, mc)
    with H.h5open(path, "r") as f:
        R = H.read_unit_cell(f)
        _same_uc(uc, R)
        L2 = H.read_lattice(f, seed=1)
        for k in ("S", "bc", "size", "shape"):
            assert getattr(lat, k) == getattr(L2, k), k
        assert np.array_equal(lat.site_positions, L2.site_positions)
        t1, t2 = lat.tables(), L2.tables()
        for k in t1:
            assert np.array_equal(t1[k], t2[k]), k
        attrs = H.attributes_to_dict(f)
        assert attrs["t_thermalization"] == 10 ** 4 and attrs["report_interval"] == 0
    os.unlink(path)


def test_on_disk_names_and_shapes(tmp_path):
    # the names util/load.py reads, with Julia's reversed dimensions
    uc = W.kgg_test_cell()
    lat = csmc.Lattice((4, 4), uc, 1.0, seed=2)
    out = str(tmp_path) + "/"
    mc = csmc.MonteCarlo(0.5, lat, {"t_thermalization": 100}, outpath=out, inparams={"K": -1.0, "tag": "x"})
    assert os.path.isfile(out + "configuration.h5.params")
    assert os.path.isfile(out + "configuration_0.h5")
    with H.h5open(out + "configuration.h5.params") as f:
        assert set(f.keys()) == {"unit_cell", "lattice"}
        assert f.get_attr("K") == -1.0 and f.get_attr("tag") == "x"
        with f.group("lattice") as l:
            assert list(l.read("size")) == [4, 4] and l.read("S") == 1.0 and l.read("bc") == "periodic"
        with f.group("unit_cell") as g:
            assert g.read("lattice_vectors").shape == (2, 2)
            assert g.read("basis").shape == (2, 2)
            with g.group("bilinear") as b:
                assert sorted(b.keys()) == ["(1,2),(0, -1)", "(1,2),(0, 0)", "(1,2),(1, -1)"]
                # InteractionMatrix written as a Julia matrix: transposed on disk
                Jx = W.kitaev_matrices(0.0, -1.0, 0.2, -0.02)[0]
                assert np.array_equal(b.read("(1,2),(0, -1)"), Jx.T)
    with H.h5open(out + "configuration_0.h5") as f:
        assert f.get_attr("T") == 0.5
        assert f.get_attr("paramsfile").endswith("configuration.h5.params")
        assert f.read("spins").shape == (32, 3)
        assert f.read("site_positions").shape == (32, 2)
        assert np.array_equal(f.read("spins"), lat._spins)
    # checkpoint overwrite and reload
    mc.lattice._spins[:] = np.tile([0.0, 0.0, 1.0], (32, 1))
    H.write_MC_checkpoint(mc)
    lat2 = csmc.Lattice((4, 4), W.kgg_test_cell(), 1.0, seed=3)
    H.read_spin_configuration(lat2, out + "configuration_0.h5")
    assert np.array_equal(lat2._spins, mc.lattice._spins)


def test_final_observables_layout(tmp_path):
    uc = csmc.Square()
    lat = csmc.Lattice((2, 2), uc, 1.0, seed=2)
    out = str(tmp_path) + "/"
    mc = csmc.MonteCarlo(0.5, lat, {}, outpath=out)
    mc.observables = dict(specific_heat=2.0, specific_heat_err=0.5, susceptibility=4.0,
                          susceptibility_err=1.0, magnetization=3.0, magnetization_err=0.1,
                          energy=-1.0, energy_err=0.01)
    H.write_final_observables(mc)
    H.write_final_observables(mc)  # second call overwrites the group
    with H.h5open(mc.outpath) as f:
        with f.group("observables") as g:
            assert sorted(g.keys()) == sorted(["specific_heat", "specific_heat_err", "susceptibility",
                                               "susceptibility_err", "magnetization",
                                               "magnetization_err", "energy", "energy_err"])
            assert g.read("specific_heat_err") == 0.25  # relative, hdf5.jl:221
            assert g.read("susceptibility_err") == 0.25
            assert g.read("energy_err") == 0.01        # absolute, hdf5.jl:227
