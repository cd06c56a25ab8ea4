"""Equal-time spin correlations, host side: the oracle's restatement of
compute_equal_time_correlations (spin_correlations.jl:6-43) pinned to
closed forms and to an independent method (an FFT of the spins on the cell
grid), and the reciprocal.jl momentum helpers checked against their
definitions."""
import numpy as np
import pytest

import oracle.oracle as orc
from csmc.reciprocal import get_allowed_wavevectors, get_k_path, reciprocal
from csmc import workloads as W
from csmc.unitcell import Honeycomb, Square, addBasisSite


def square_positions(L):
    # site order (basis, x1, x2), x2 fastest (lattice.jl:29-33)
    x1, x2 = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    return np.vstack([x1.ravel(), x2.ravel()]).astype(float), x1.ravel(), x2.ravel()


def grid_momenta(L):
    # the L x L distinct allowed momenta 2 pi m / L of the unit square lattice
    m1, m2 = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    return 2 * np.pi * np.vstack([m1.ravel(), m2.ravel()]) / L


def test_ferromagnet_and_neel_closed_forms():
    L, S = 6, 1.5
    pos, x1, x2 = square_positions(L)
    N = L * L
    ks = grid_momenta(L)
    fm = np.zeros((N, 3))
    fm[:, 2] = S
    Suv = orc.equal_time_correlations(pos, fm, ks)
    assert Suv.shape == (9, N)
    assert Suv[8, 0] == pytest.approx(N * S * S, rel=1e-13)
    assert np.max(np.abs(Suv[8, 1:])) < 1e-12
    assert np.max(np.abs(np.delete(Suv, 8, axis=0))) < 1e-12
    neel = np.zeros((N, 3))
    neel[:, 2] = S * (-1.0) ** (x1 + x2)
    Suv = orc.equal_time_correlations(pos, neel, np.array([[np.pi], [np.pi]]))
    assert Suv[8, 0] == pytest.approx(N * S * S, rel=1e-13)


def test_parseval_symmetry_and_fft():
    rng = np.random.default_rng(3)
    L = 8
    pos, x1, x2 = square_positions(L)
    N = L * L
    v = rng.normal(size=(N, 3))
    spins = 0.5 * v / np.linalg.norm(v, axis=1, keepdims=True)
    ks = grid_momenta(L)
    Suv = orc.equal_time_correlations(pos, spins, ks)
    # Parseval: sum over the N distinct momenta of S^aa = sum_i s_ia^2
    for a in range(3):
        assert Suv[4 * a].sum() == pytest.approx((spins[:, a] ** 2).sum(), rel=1e-12)
    # S^ab = S^ba
    for a in range(3):
        for b in range(3):
            assert np.allclose(Suv[3 * a + b], Suv[3 * b + a], rtol=1e-12, atol=1e-14)
    # S(k) = S(-k)
    assert np.allclose(orc.equal_time_correlations(pos, spins, -ks), Suv, rtol=1e-12, atol=1e-13)
    # an FFT of the spins on the cell grid gives the same s_a(k)
    grid = spins.reshape(L, L, 3)
    F = np.fft.fft2(grid, axes=(0, 1)).reshape(N, 3)  # sum_x exp(-2 pi i m.x / L) s(x)
    ref = np.einsum("ka,kb->abk", F, np.conj(F)).real.reshape(9, N) / N
    assert np.allclose(Suv, ref, rtol=1e-11, atol=1e-12)


def test_planted_sign_error_is_caught():
    """a wrong sign of the phase leaves Parseval intact but breaks the FFT
    comparison for an off-diagonal correlation of a chiral texture"""
    L = 8
    pos, x1, x2 = square_positions(L)
    N = L * L
    th = 2 * np.pi * x1 / L
    spins = np.stack([np.cos(th), np.sin(th), np.zeros(N)], axis=1) * 0.5
    ks = grid_momenta(L)
    good = orc.equal_time_correlations(pos, spins, ks)
    grid = spins.reshape(L, L, 3)
    F = np.fft.fft2(grid, axes=(0, 1)).reshape(N, 3)
    ref = np.einsum("ka,kb->abk", F, np.conj(F)).real.reshape(9, N) / N
    assert np.allclose(good, ref, atol=1e-12)
    # an imaginary-part-only off-diagonal term (Im s_x conj(s_y)) is what a
    # dropped minus sign would leak into S^xy: it is non-zero here
    leak = np.einsum("ka,kb->abk", F, np.conj(F)).imag.reshape(9, N) / N
    assert np.max(np.abs(leak[1])) > 1e-3


def test_reciprocal_vectors():
    for uc in (Square(), Honeycomb(), W.pyrochlore_mixed()):
        a = [np.asarray(v, dtype=float) for v in uc.lattice_vectors]
        b = reciprocal(*a)
        G = np.array([[np.dot(bi, aj) for aj in a] for bi in b])
        assert np.allclose(G, 2 * np.pi * np.eye(len(a)), atol=1e-12)


def test_allowed_wavevectors_order_and_count():
    uc = Square()
    addBasisSite(uc, [0.0, 0.0])
    ks = get_allowed_wavevectors(uc, (4, 3))
    assert ks.shape == (2, 5 * 4)
    # the first dimension varies fastest (Iterators.product)
    assert np.allclose(ks[:, 1], [2 * np.pi / 4, 0.0])
    assert np.allclose(ks[:, 5], [0.0, 2 * np.pi / 3])
    ks = get_allowed_wavevectors(uc, (4, 4), min=-1, max=1)
    assert ks.shape == (2, 81)
    assert np.allclose(ks.min(axis=1), [-2 * np.pi, -2 * np.pi])


def test_k_path_along_a_direction_and_through_points():
    uc = Square()
    addBasisSite(uc, [0.0, 0.0])
    line = get_k_path(uc, [1.0, 0.0], (4, 4))
    assert line.shape[1] > 0 and np.allclose(line[1], 0.0)
    hsp = {"G": np.array([0.0, 0.0]), "X": np.array([np.pi, 0.0]), "M": np.array([np.pi, np.pi])}
    counts, path = get_k_path(uc, hsp, ["G", "X", "M"], (8, 8))
    assert counts[0] == 0 and counts[-1] == path.shape[1] - 1
    assert np.allclose(path[:, 0], [0, 0]) and np.allclose(path[:, -1], [np.pi, np.pi])
    steps = np.linalg.norm(np.diff(path, axis=1), axis=0)
    assert np.allclose(steps, 2 * np.pi / 8)


def test_structure_factor_average_over_files(tmp_path):
    """compute_equal_time_structure_factor (spin_correlations.jl:116-145):
    the mean of every file's SSF (files in name order), written with the
    first file's momenta to dest; the 9 x n_k matrices are stored as Julia
    lays them out (row-major n_k x 9)"""
    from csmc import hdf5io as H
    from csmc.spin_correlations import compute_equal_time_structure_factor
    rng = np.random.default_rng(1)
    src = tmp_path / "ics"
    src.mkdir()
    ks = rng.normal(size=(2, 7))
    mats = []
    for i in range(4):
        S = rng.normal(size=(9, 7))
        mats.append(S)
        with H.h5open(str(src / f"IC_{i}.h5"), "w") as f:
            with f.group("spin_correlations", create=True) as g:
                g.write("SSF", np.ascontiguousarray(S.T))
                g.write("SSF_momentum", np.ascontiguousarray(ks.T))
    dest = str(tmp_path / "dest.h5")
    with H.h5open(dest, "w") as f:
        f.write("spins", np.zeros((3, 3)))
    mean, kout = compute_equal_time_structure_factor(str(src), dest)
    ref = (mats[0] + mats[1] + mats[2] + mats[3]) / 4
    assert np.allclose(mean, ref, rtol=1e-15, atol=0)
    assert np.array_equal(kout, ks)
    with H.h5open(dest, "r") as f:
        assert np.allclose(f.read("spin_correlations/SSF"), ref.T, rtol=1e-15, atol=0)
        assert np.array_equal(f.read("spin_correlations/SSF_momentum"), ks.T)
    # a second run overwrites the keys in place
    compute_equal_time_structure_factor(str(src), dest)
    with H.h5open(dest, "r") as f:
        assert np.allclose(f.read("spin_correlations/SSF"), ref.T, rtol=1e-15, atol=0)


def test_montecarlo_corr_arguments():
    """MonteCarlo(...; corr, ks) keeps the momenta as the D x n_k matrix and
    warns as the reference does (monte_carlo.jl:64-68)"""
    from csmc import Lattice, MonteCarlo
    lat = Lattice((4, 4), W.square_heisenberg(-1.0, (0, 0, 0.1)), 1.0, seed=1)
    ks = np.zeros((2, 5))
    mc = MonteCarlo(1.0, lat, {"t_thermalization": 10}, corr=True, ks=ks)
    assert mc.corr and mc.momentum_vectors.shape == (2, 5)
    with pytest.warns(UserWarning, match="No momentum vectors"):
        MonteCarlo(1.0, lat, {}, corr=True)
    with pytest.warns(UserWarning, match="not requested"):
        MonteCarlo(1.0, lat, {}, corr=False, ks=ks)
