"""Pins the CPU oracle (oracle/csmc_oracle.c) before it is trusted as the
checker: the reference's own known-answer tests, closed forms, brute force
on tiny lattices and the Philox known-answer vectors.  No GPU."""
import math

import numpy as np
import pytest

import oracle.oracle as orc
from csmc import workloads as W
from csmc.unitcell import (Honeycomb, Pyrochlore, Square, Triangular, addBilinear, addCubic,
                           addOnSite, addQuartic, addZeemanCoupling)
from conftest import random_spins


# -- test/latticetests.jl KATs ------------------------------------------------

def test_zeeman_kat():
    # test/latticetests.jl:10-19: E == -1.0 and H == (-1,0,0) exactly
    uc = Square()
    uc.basis.append(np.zeros(2))
    addZeemanCoupling(uc, 1, [1.0, 0.0, 0.0])
    T = orc.Tables(uc, (1, 1))
    s = np.array([[1.0, 0.0, 0.0]])
    assert orc.total_energy(T, s) == -1.0
    assert tuple(orc.local_field(T, s, 1)) == (-1.0, -0.0, -0.0)


def test_bilinear_kat():
    # test/latticetests.jl:21-31: FM on 2x2 square, E/N == -2.0 exactly
    uc = Square()
    uc.basis.append(np.zeros(2))
    for d in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        addBilinear(uc, 1, 1, -np.eye(3), d)
    T = orc.Tables(uc, (2, 2))
    s = np.tile([1.0, 0.0, 0.0], (4, 1))
    assert orc.total_energy(T, s) / 4 == -2.0


def _cells():
    """unit cells covering every table path (same-basis, cross-basis,
    cubic/quartic role permutations, onsite, open boundaries)"""
    out = []
    uc = W.square_heisenberg()
    uc.basis.append(np.zeros(2))
    out.append(("square", uc, (4, 3)))
    out.append(("honeycomb_kgg", W.kgg_test_cell(), (3, 4)))
    out.append(("tri_j1j2", _with_basis(W.triangular_j1j2(1.0, 0.13)), (4, 5)))
    out.append(("pyro_mixed", W.pyrochlore_mixed(), (2, 2, 2)))
    uc = Honeycomb()
    rng = np.random.default_rng(3)
    addCubic(uc, 2, 1, 2, rng.normal(size=(3, 3, 3)), (1, 0), (0, -1))
    addQuartic(uc, 1, 2, 2, 1, rng.normal(size=(3, 3, 3, 3)), (0, 0), (1, -1), (-1, 1))
    addBilinear(uc, 2, 1, rng.normal(size=(3, 3)), (1, 1))
    addOnSite(uc, 1, rng.normal(size=(3, 3)))
    addOnSite(uc, 2, rng.normal(size=(3, 3)))
    addZeemanCoupling(uc, 2, rng.normal(size=3))
    addZeemanCoupling(uc, 1, rng.normal(size=3))
    out.append(("honeycomb_random_terms", uc, (3, 3)))
    return out


def _with_basis(uc):
    if not uc.basis:
        uc.basis.append(np.zeros(uc.D))
    return uc


@pytest.mark.parametrize("periodic", [True, False])
def test_literal_findfirst_equals_arithmetic(periodic):
    # lattice.jl:196 scans the sorted tuple list; the oracle's arithmetic
    # lookup must give the same tables
    for name, uc, shape in _cells():
        a = orc.Tables(uc, shape, periodic, search=0)
        b = orc.Tables(uc, shape, periodic, search=1)
        for k in ("bsite", "bmat", "csite", "ctens", "qsite", "qtens", "field", "onsite", "pos"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), (name, k)


def test_neighbour_geometry_brute_force():
    # every bilinear entry of site i points at the site whose position is
    # pos_i + (basis_bj - basis_bi) + offset . a  (mod the periodic box)
    for name, uc, shape in _cells():
        T = orc.Tables(uc, shape, True)
        D = len(uc.lattice_vectors)
        A = np.stack(uc.lattice_vectors)
        box = np.array(shape)
        cell = np.linalg.inv(A.T)
        nb = len(uc.basis)
        ncell = int(np.prod(shape))
        for i in range(T.N):
            bi = i // ncell
            for t, (b1, b2, M, off) in enumerate(uc.bilinear):
                j = T.bsite[i, t]
                if bi + 1 not in (b1, b2):
                    assert j == 0
                    continue
                assert j > 0
                if b1 == b2 or b1 == bi + 1:
                    bj, o = b2 if b1 != b2 else bi + 1, np.array(off)
                else:
                    bj, o = b1, -np.array(off)
                expect = T.pos[i] - uc.basis[bi] + uc.basis[bj - 1] + o @ A
                d = expect - T.pos[j - 1]
                frac = cell @ d  # lattice coordinates of the difference
                assert np.allclose(frac - box * np.round(frac / box), 0, atol=1e-9), (name, i, t)


def test_pyrochlore_bonds_are_nearest_neighbours():
    uc = Pyrochlore()
    A = np.stack(uc.lattice_vectors)
    for b1, b2, o in W.PYRO_BONDS:
        d = uc.basis[b2 - 1] + np.array(o) @ A - uc.basis[b1 - 1]
        assert abs(np.linalg.norm(d) - math.sqrt(2) / 4) < 1e-12


def test_q2_field_quirks():
    # lattice.jl:128-140 indexes the call list by basis number
    uc = Honeycomb()
    addZeemanCoupling(uc, 2, [1.0, 2.0, 3.0])  # only basis 2 of 2: BoundsError
    with pytest.raises(orc.OracleError) as e:
        orc.Tables(uc, (2, 2))
    assert e.value.rc == 1
    uc = Honeycomb()
    addZeemanCoupling(uc, 2, [1.0, 2.0, 3.0])
    addZeemanCoupling(uc, 1, [4.0, 5.0, 6.0])  # both, out of order: correct
    T = orc.Tables(uc, (2, 2))
    assert np.array_equal(T.field[0], [4, 5, 6]) and np.array_equal(T.field[-1], [1, 2, 3])


def test_ground_state_closed_forms():
    # config 1 FM ground state along +z: E/N = -2 S^2 |J| - h S = -2.1
    uc = _with_basis(W.square_heisenberg(-1.0, (0, 0, 0.1)))
    T = orc.Tables(uc, (16, 16))
    s = np.tile([0.0, 0.0, 1.0], (256, 1))
    assert abs(orc.total_energy(T, s) / 256 + 2.1) < 1e-14
    # honeycomb K=-2, J=1 at S=1/2: the Klein-dual FM ground state on a 2x2
    # supercell, E/N = -(3/2) S^2 = -0.375
    uc = W.honeycomb_kitaev(1.0, -2.0)
    T = orc.Tables(uc, (4, 4))
    s = _klein_state(T, 0.5, uc)
    assert abs(orc.total_energy(T, s) / T.N + 0.375) < 1e-13


def _klein_state(T, S, uc):
    """Klein dual of the FM state along +x for the honeycomb J-K model: every
    site of the periodic 2x2 cell (8 sites) gets the identity or a pi
    rotation about x, y or z; the assignment is found by enumerating all
    4^8 of them, then tiled with period 2."""
    import itertools
    ncell = T.N // 2
    L1 = L2 = int(round(math.sqrt(ncell)))
    rots = np.array([np.diag(v) for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1])], float)
    base = np.array([1.0, 0.0, 0.0]) * S
    cand = rots @ base  # (4, 3)
    Ttiny = orc.Tables(uc, (2, 2))
    best = None
    for choice in itertools.product(range(4), repeat=8):
        s = np.ascontiguousarray(cand[list(choice)])
        e = orc.total_energy(Ttiny, s)
        if best is None or e < best[0] - 1e-12:
            best = (e, choice)
    assert abs(best[0] / 8 + 0.375) < 1e-14
    choice = best[1]
    s = np.zeros((T.N, 3))
    for b in range(2):
        for x1 in range(L1):
            for x2 in range(L2):
                s[b * ncell + x1 * L2 + x2] = cand[choice[b * 4 + (x1 % 2) * 2 + x2 % 2]]
    return s


def test_cubic_quartic_closed_forms():
    # chirality of a right-handed orthonormal triad is +1; biquadratic term
    # B (s.s')^2 with the repeated-site quartic (Q3: total_energy reports
    # B/2 per bond, energy() samples B)
    uc = Square()
    uc.basis.append(np.zeros(2))
    addCubic(uc, 1, 1, 1, W.levi_civita(), (1, 0), (0, 1))
    T = orc.Tables(uc, (3, 3))
    s = np.zeros((9, 3))
    s[:] = [1, 0, 0]
    s[3] = [0, 1, 0]  # (x1=1, x2=0) = +a1 of site 0
    s[1] = [0, 0, 1]  # (0, 1) = +a2
    assert abs(orc.site_energy(T, s, 1) - 1.0) < 1e-15
    uc = Square()
    uc.basis.append(np.zeros(2))
    B = -0.1
    addQuartic(uc, 1, 1, 1, 1, B * np.einsum("ab,cd->abcd", np.eye(3), np.eye(3)), (1, 0), (0, 0),
               (1, 0))
    T = orc.Tables(uc, (3, 1))
    rng = np.random.default_rng(1)
    s = random_spins(3, 1.0, rng)
    bonds = sum(float(np.dot(s[i], s[(i + 1) % 3])) ** 2 for i in range(3))
    # each site holds the entry to its +x neighbour: E = sum_i B (s_i.s_i+1)^2 / 4
    assert abs(orc.total_energy(T, s) - B * bonds / 4) < 1e-14
    assert abs(orc.site_energy(T, s, 1) - B * float(np.dot(s[0], s[1])) ** 2) < 1e-15


def test_local_field_is_energy_gradient():
    # for symmetric tables (reverse bonds carry J^T, symmetric onsite,
    # distinct sites per term) dE/ds_i == H_i; central differences
    rng = np.random.default_rng(7)
    uc = W.honeycomb_kitaev(1.0, -2.0, 0.3, -0.1, h=[0.1, 0.2, -0.3])
    O = rng.normal(size=(3, 3))
    addOnSite(uc, 1, O + O.T)
    addOnSite(uc, 2, O @ O.T)
    T = orc.Tables(uc, (3, 3))
    s = random_spins(T.N, 1.0, rng)
    eps = 1e-6
    for i in [0, 4, 11, 17]:
        H = orc.local_field(T, s, i + 1)
        g = np.zeros(3)
        for a in range(3):
            sp, sm = s.copy(), s.copy()
            sp[i, a] += eps
            sm[i, a] -= eps
            g[a] = (orc.total_energy(T, sp) - orc.total_energy(T, sm)) / (2 * eps)
        assert np.allclose(g, H, rtol=1e-7, atol=1e-8)


def test_overrelaxation_invariants():
    rng = np.random.default_rng(11)
    uc = W.honeycomb_kitaev(1.0, -2.0, h=[0.0, 0.1, 0.2])
    T = orc.Tables(uc, (4, 4))
    s = random_spins(T.N, 0.5, rng)
    E0 = orc.total_energy(T, s)
    H0 = orc.local_fields(T, s)
    dots0 = np.einsum("ij,ij->i", s, H0)
    sub = np.arange(1, T.N // 2 + 1)  # sublattice 1: an independent set
    s1 = s.copy()
    orc.overrelax(T, s1, sub)
    assert np.allclose(np.linalg.norm(s1, axis=1), 0.5, atol=1e-15)
    dots1 = np.einsum("ij,ij->i", s1, orc.local_fields(T, s1))
    assert np.allclose(dots0[: T.N // 2], dots1[: T.N // 2], atol=1e-14)
    assert abs(orc.total_energy(T, s1) - E0) < 1e-12  # bilinear+Zeeman: E conserved
    s2 = s1.copy()
    orc.overrelax(T, s2, sub)  # involution on an independent set
    assert np.allclose(s2, s, atol=1e-14)


def test_random_spin_formula():
    # lattice.jl:306-311
    o = orc.random_spin(2.0, 0.25, 0.75)
    assert np.allclose(o, [2.0 * math.sqrt(0.75) * math.cos(math.pi / 2),
                           2.0 * math.sqrt(0.75) * math.sin(math.pi / 2), 1.0], atol=1e-15)


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert list(orc.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    ff = 0xFFFFFFFF
    assert list(orc.philox([ff] * 4, [ff, ff])) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                           [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_binner_matches_direct_blocking():
    rng = np.random.default_rng(5)
    x = np.cumsum(rng.normal(size=4096)) * 0.01 + rng.normal(size=4096)
    b = orc.Binner()
    for v in x:
        b.push(v)
    assert abs(b.means()[0] - x.mean()) < 1e-12
    for lvl in range(6):
        blocks = x[: (len(x) >> lvl) << lvl].reshape(-1, 1 << lvl).mean(axis=1)
        v_direct = blocks.var(ddof=1)
        assert abs(b.var([1.0, 0.0], lvl) - v_direct) < 1e-10 * max(1.0, v_direct)
    # 4096 samples -> level with >= 32 bins is 7 (4096/128 = 32)
    assert b.reliable_level() == 7


def test_single_spin_langevin():
    # one spin in a field (no bonds): <s.h^> = S L(beta h S), L = coth x - 1/x
    uc = Square()
    uc.basis.append(np.zeros(2))
    addZeemanCoupling(uc, 1, [0.0, 0.0, 0.7])
    T = orc.Tables(uc, (1, 1))
    s = np.array([[0.0, 0.0, 1.0]])
    temp = 0.5
    zs = []
    for sw in range(40000):
        orc.metropolis_random(T, s, 1.0, temp, 0, 60.0, 99, 0, sw)
        zs.append(s[0, 2])
    x = 0.7 / temp
    exact = 1.0 / math.tanh(x) - 1.0 / x
    zs = np.array(zs[1000:])
    err = zs.reshape(-1, 100).mean(axis=1).std(ddof=1) / math.sqrt(len(zs) / 100)
    assert abs(zs.mean() - exact) < 4 * err


def test_exchange_two_level_detailed_balance():
    # two replicas, energies fixed: the fraction of accepted swaps from
    # (E0 at T0, E1 at T1) is min(1, exp((1/T1-1/T0)(E1-E0)))
    temps = np.array([0.5, 1.0])
    n, acc = 20000, 0
    for xc in range(0, 2 * n, 2):
        E = np.array([-1.0, 0.0])
        perm = np.array([0, 1], dtype=np.intc)
        acc += orc.pt_exchange(temps, E, perm, 4242, xc)
    p = min(1.0, math.exp((1 / 1.0 - 1 / 0.5) * (0.0 - -1.0)))
    assert abs(acc / n - p) < 4 * math.sqrt(p * (1 - p) / n)


# -- cubic / quartic role permutations (lattice.jl:209-286) ---------------------
# The reference tests hold no cubic or quartic case, so these pin the oracle's
# role branches against what the mathematics fixes: a term-by-term energy sum
# written independently of the per-site tables, and the gradient of that
# energy.

def _site_index(shape, b, x):
    """0-based site index of basis b (1-based) in cell x (0-based), the
    lexicographic (basis, x1, .., xD) order of lattice.jl:29-33"""
    idx = b - 1
    for d, L in enumerate(shape):
        idx = idx * L + (x[d] % L)
    return idx


def _term_energy(uc, shape, s):
    """Physical energy of the cubic and quartic terms of uc on a periodic
    lattice: every term instance counted once, sum over cells x of
    C[a,b,c] s_b1(x)_a s_b2(x+o2)_b s_b3(x+o3)_c (and the quartic alike)."""
    import itertools
    E = 0.0
    for x in itertools.product(*[range(L) for L in shape]):
        x = np.array(x)
        for b1, b2, b3, C, o2, o3 in uc.cubic:
            i, j, k = (_site_index(shape, b1, x), _site_index(shape, b2, x + np.array(o2)),
                       _site_index(shape, b3, x + np.array(o3)))
            E += np.einsum("abc,a,b,c->", C, s[i], s[j], s[k])
        for b1, b2, b3, b4, R, o2, o3, o4 in uc.quartic:
            i, j = _site_index(shape, b1, x), _site_index(shape, b2, x + np.array(o2))
            k, m = _site_index(shape, b3, x + np.array(o3)), _site_index(shape, b4, x + np.array(o4))
            E += np.einsum("abcd,a,b,c,d->", R, s[i], s[j], s[k], s[m])
    return E


def _generic_pyrochlore_cell(seed=21):
    """pyrochlore with non-symmetric cubic tensors on faces of three distinct
    bases (each basis in every role position somewhere) and generic 4-site
    quartic tensors, with offsets that are not all zero, plus a field on every
    basis (in basis order, the safe case of lattice.jl:128-140)"""
    rng = np.random.default_rng(seed)
    uc = Pyrochlore()
    addCubic(uc, 1, 2, 3, rng.normal(size=(3, 3, 3)), (0, 0, 0), (0, 0, 0))
    addCubic(uc, 4, 1, 2, rng.normal(size=(3, 3, 3)), (1, 0, 0), (0, 1, -1))
    addCubic(uc, 3, 4, 1, rng.normal(size=(3, 3, 3)), (-1, 1, 0), (0, 0, 1))
    addQuartic(uc, 1, 2, 3, 4, rng.normal(size=(3, 3, 3, 3)), (0, 0, 0), (0, 0, 0), (0, 0, 0))
    addQuartic(uc, 3, 1, 4, 2, rng.normal(size=(3, 3, 3, 3)), (1, 0, 0), (0, -1, 1), (1, 1, 0))
    addQuartic(uc, 2, 4, 1, 3, rng.normal(size=(3, 3, 3, 3)), (0, 1, 0), (-1, 0, 1), (0, 0, -1))
    for b in range(1, 5):
        addZeemanCoupling(uc, b, rng.normal(size=3))
    return uc


def test_pyrochlore_cubic_quartic_roles_pinned():
    # total_energy (hamiltonian.jl:70-132, weights 1/3 and 1/4 at :131) equals
    # the term-by-term sum when every term has distinct sites, and
    # get_local_field (hamiltonian.jl:3-67) is its gradient: both fail if a
    # role permutation (lattice.jl:215-227,251-272) or a role offset
    # (lattice.jl:212-224,244-268) were misread.
    uc = _generic_pyrochlore_cell()
    shape = (3, 3, 3)
    T = orc.Tables(uc, shape)
    rng = np.random.default_rng(4)
    s = random_spins(T.N, 1.0, rng)
    zeeman = sum(-np.dot(uc.field[b][1], s[i]) for b in range(4)
                 for i in range(b * 27, (b + 1) * 27))
    E = orc.total_energy(T, s)
    assert abs(E - (_term_energy(uc, shape, s) + zeeman)) < 1e-11 * max(1.0, abs(E))
    eps = 1e-6
    worst = 0.0
    for i in (0, 13, 27, 40, 54, 67, 81, 107):  # sites of every basis
        H = orc.local_field(T, s, i + 1)
        g = np.zeros(3)
        for a in range(3):
            sp, sm = s.copy(), s.copy()
            sp[i, a] += eps
            sm[i, a] -= eps
            g[a] = (orc.total_energy(T, sp) - orc.total_energy(T, sm)) / (2 * eps)
        worst = max(worst, float(np.max(np.abs(g - H))))
    assert worst <= 1e-7, worst


def test_planted_role_permutation_mistake_is_caught():
    # the same check run on a deliberately wrong contraction: using the
    # unpermuted tensor for a role-2 site gives a field that is not the
    # gradient, so the check above can fail
    uc = Pyrochlore()
    rng = np.random.default_rng(3)
    C = rng.normal(size=(3, 3, 3))
    addCubic(uc, 1, 2, 3, C, (0, 0, 0), (1, 0, 0))
    shape = (2, 2, 2)
    T = orc.Tables(uc, shape)
    s = random_spins(T.N, 1.0, rng)
    i = _site_index(shape, 2, (1, 0, 1))      # a basis-2 site: role 2 of its term
    x = np.array((1, 0, 1))
    i1, i3 = _site_index(shape, 1, x), _site_index(shape, 3, x + np.array((1, 0, 0)))
    right = np.einsum("abc,a,c->b", C, s[i1], s[i3])
    wrong = np.einsum("abc,b,c->a", C, s[i1], s[i3])
    H = orc.local_field(T, s, i + 1)
    assert np.allclose(H, right, atol=1e-13)
    assert not np.allclose(H, wrong, atol=1e-3)


def test_repeated_basis_term_first_matching_role():
    # Reference quirk: the role chain of lattice.jl:212-224 / :244-268 is an
    # if/elseif, so a site takes only its FIRST matching role.  In a term that
    # repeats a basis (distinct sites, e.g. addCubic!(2,1,1,...)) the later
    # roles of that basis get no entry.  Then total_energy counts the term
    # (number of distinct roles hit)/3 (or /4) times, and get_local_field is
    # the derivative with respect to the first matching role only.
    uc = Honeycomb()
    rng = np.random.default_rng(8)
    C = rng.normal(size=(3, 3, 3))
    addCubic(uc, 2, 1, 1, C, (1, 0), (0, 1))              # basis 1 matches role 2, never 3
    shape = (3, 3)
    T = orc.Tables(uc, shape)
    s = random_spins(T.N, 1.0, rng)
    Et = _term_energy(uc, shape, s)
    assert abs(orc.total_energy(T, s) - 2.0 / 3.0 * Et) < 1e-12
    H = np.zeros((T.N, 3))
    import itertools
    for x in itertools.product(range(3), range(3)):
        x = np.array(x)
        i = _site_index(shape, 2, x)
        j = _site_index(shape, 1, x + (1, 0))
        k = _site_index(shape, 1, x + (0, 1))
        H[i] += np.einsum("abc,b,c->a", C, s[j], s[k])   # role 1
        H[j] += np.einsum("abc,a,c->b", C, s[i], s[k])   # role 2 (first match of basis 1)
    assert np.allclose(orc.local_fields(T, s), H, atol=1e-12)
    # so the field is not the gradient of either energy
    eps = 1e-6
    p = _site_index(shape, 1, (0, 0))
    g = np.zeros(3)
    for a in range(3):
        sp, sm = s.copy(), s.copy()
        sp[p, a] += eps
        sm[p, a] -= eps
        g[a] = (orc.total_energy(T, sp) - orc.total_energy(T, sm)) / (2 * eps)
    assert np.max(np.abs(g - H[p])) > 1e-2

    uc = Honeycomb()
    R = rng.normal(size=(3, 3, 3, 3))
    addQuartic(uc, 2, 1, 2, 1, R, (0, 0), (1, 1), (-1, 1))  # roles 3 and 4 never taken
    T = orc.Tables(uc, shape)
    assert abs(orc.total_energy(T, s) - 0.5 * _term_energy(uc, shape, s)) < 1e-12
    H = np.zeros((T.N, 3))
    for x in itertools.product(range(3), range(3)):
        x = np.array(x)
        i, j = _site_index(shape, 2, x), _site_index(shape, 1, x)
        k, m = _site_index(shape, 2, x + (1, 1)), _site_index(shape, 1, x + (-1, 1))
        H[i] += np.einsum("abcd,b,c,d->a", R, s[j], s[k], s[m])
        H[j] += np.einsum("abcd,a,c,d->b", R, s[i], s[k], s[m])
    assert np.allclose(orc.local_fields(T, s), H, atol=1e-12)


@pytest.mark.parametrize("sigma", [0.3, 1.0])
def test_cone_move_langevin(sigma):
    # the cone proposal s + sigma * (uniform * S), renormalised to S
    # (metropolis.jl:84-87), is symmetric, so the chain samples the Boltzmann
    # distribution: <s.h^> = S L(beta h S) for one spin in a field
    uc = Square()
    uc.basis.append(np.zeros(2))
    addZeemanCoupling(uc, 1, [0.0, 0.0, 0.7])
    T = orc.Tables(uc, (1, 1))
    s = np.array([[0.0, 0.0, 1.0]])
    temp = 0.5
    zs = np.empty(60000)
    for sw in range(len(zs)):
        orc.metropolis_random(T, s, 1.0, temp, 1, sigma, 77, 0, sw)
        zs[sw] = s[0, 2]
    x = 0.7 / temp
    exact = 1.0 / math.tanh(x) - 1.0 / x
    zs = zs[2000:]
    nb = 200
    blocks = zs[: len(zs) // nb * nb].reshape(nb, -1).mean(axis=1)
    err = blocks.std(ddof=1) / math.sqrt(nb)
    assert abs(zs.mean() - exact) < 4 * err, (zs.mean(), exact, err)


def test_cone_move_formula():
    # metropolis.jl:84-87: normalise(s + sigma * random_spin(S)) * S
    S, sigma = 1.5, 0.4
    s = np.array([0.3, -1.2, 0.8])
    s = S * s / np.linalg.norm(s)
    u1, u2 = 0.2, 0.9
    r = orc.random_spin(S, u1, u2)
    v = s + sigma * r
    assert np.allclose(orc.cone_move(S, s, sigma, u1, u2), S * v / np.linalg.norm(v), atol=1e-14)


# -- metropolis_constraint_adaptive! (metropolis.jl:155-179) --------------------

def _free_spin():
    uc = Square()
    uc.basis.append(np.zeros(2))
    return orc.Tables(uc, (1, 1))


def test_constraint_weight_zero_is_the_cone_chain():
    # weight 0: new_lambda = lambda = 0, so the acceptance is on dE alone and
    # the chain is the cone chain (metropolis_fixed_cone!) draw for draw
    uc = W.honeycomb_kitaev(1.0, -2.0, h=[0.1, 0.0, 0.2])
    T = orc.Tables(uc, (4, 4))
    rng = np.random.default_rng(2)
    s = random_spins(T.N, 0.5, rng)
    order = np.arange(1, T.N + 1)
    a, b = s.copy(), s.copy()
    sigma = 0.7
    for sw in range(3):
        na, deca, sg, lam = orc.metropolis_constraint_coloured(
            T, a, order, 0.5, 0.4, sigma, 60.0, 11, 0, sw, T.N // 2, [[0, 0, 1], [0, 0, 1]],
            np.zeros((2, 9)), 3.0, 0.0, 0.0)
        nb, decb, _ = orc.metropolis_coloured(T, b, order, 0.5, 0.4, 1, sigma, 11, 0, sw)
        assert lam == 0.0 and na == nb and np.array_equal(deca, decb)
        assert np.array_equal(a, b)
        # the width law of metropolis.jl:176-177
        rate = na / T.N
        assert sg == (sigma * 0.5 / (1 - rate) if rate < 0.5 else 60.0)
        sigma = sg


def test_constraint_multiplier_recurrence():
    # one free spin (dE = 0), c = s_z - 2 (always negative for S = 1):
    # new_lambda = lambda - w c > lambda, and x = dE - new_lambda c > 0.
    # At T = 1e300 every proposal is accepted, so lambda = -w sum_k c_k with
    # c_k taken with the NEW spin in place (metropolis.jl:164-165); at
    # T = 1e-300 none is, and lambda stays 0 (it moves only on acceptance,
    # :172).  The width law gives sigma0 at rate 1 and sigma / 2 at rate 0.
    T = _free_spin()
    w, c0 = 0.3, 2.0
    s = np.array([[0.0, 0.0, 1.0]])
    lam, expect = 0.0, 0.0
    for sw in range(6):
        n, dec, sg, lam = orc.metropolis_constraint_coloured(T, s, [1], 1.0, 1e300, 0.9, 60.0, 5, 0, sw, 1,
                                                             [[0, 0, 1]], np.zeros((1, 9)), c0, w, lam)
        assert n == 1 and dec[0] == 1 and sg == 60.0
        expect -= w * (s[0, 2] - c0)
        assert abs(lam - expect) < 1e-14
    s0 = s.copy()
    n, dec, sg, lam0 = orc.metropolis_constraint_coloured(T, s, [1], 1.0, 1e-300, 0.9, 60.0, 5, 0, 99, 1,
                                                          [[0, 0, 1]], np.zeros((1, 9)), c0, w, 0.0)
    assert n == 0 and lam0 == 0.0 and np.array_equal(s, s0) and sg == 0.45


def test_constraint_quadratic_term():
    # the quadratic part s.Q s enters c like the reference's closure would:
    # with Q = I the term is S^2 per site, a constant, so c is the same for
    # every proposal and the chain equals the Q = 0 chain with c0 shifted
    uc = W.honeycomb_kitaev(1.0, -2.0)
    T = orc.Tables(uc, (3, 3))
    rng = np.random.default_rng(6)
    s = random_spins(T.N, 1.0, rng)
    order = np.arange(1, T.N + 1)
    a, b = s.copy(), s.copy()
    eye = np.tile(np.eye(3).reshape(9), (2, 1))
    ra = orc.metropolis_constraint_coloured(T, a, order, 1.0, 0.3, 0.5, 60.0, 3, 1, 0, 9,
                                            [[0.2, 0, 1], [0, -0.3, 1]], eye, 5.0 + T.N, 0.2, 0.1)
    rb = orc.metropolis_constraint_coloured(T, b, order, 1.0, 0.3, 0.5, 60.0, 3, 1, 0, 9,
                                            [[0.2, 0, 1], [0, -0.3, 1]], np.zeros((2, 9)), 5.0, 0.2, 0.1)
    assert ra[0] == rb[0] and np.array_equal(ra[1], rb[1])
    assert abs(ra[3] - rb[3]) < 1e-9 * max(1.0, abs(rb[3]))
    assert np.allclose(a, b, atol=1e-12)
