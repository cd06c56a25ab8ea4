"""The adjoint of total_energy (hamiltonian.jl:70-132) on the GPU: dE/dS and
dE/dJ against central finite differences of the oracle's FP64 energy
(step 1e-6), including asymmetric same-basis bonds (only one direction added)
and repeated-site quartic terms, where dE/dS is not the local field; and
projected gradient descent against deterministic updates."""
import copy

import numpy as np
import pytest

import csmc
import oracle.oracle as orc
from csmc import workloads as W
from csmc.unitcell import (Honeycomb, Square, addBasisSite, addBilinear, addCubic, addOnSite, addQuartic,
                           addZeemanCoupling)
from conftest import random_spins

pytestmark = pytest.mark.gpu
EPS = 1e-6


def _random_cell(seed=21):
    rng = np.random.default_rng(seed)
    uc = Honeycomb()
    addCubic(uc, 2, 1, 2, rng.normal(size=(3, 3, 3)), (1, 0), (0, -1))
    addQuartic(uc, 1, 2, 2, 1, rng.normal(size=(3, 3, 3, 3)), (0, 0), (1, -1), (-1, 1))
    addBilinear(uc, 2, 1, rng.normal(size=(3, 3)), (1, 1))
    addBilinear(uc, 2, 2, rng.normal(size=(3, 3)), (0, 2))  # one direction only
    addOnSite(uc, 1, rng.normal(size=(3, 3)))
    addOnSite(uc, 2, rng.normal(size=(3, 3)))
    addZeemanCoupling(uc, 2, rng.normal(size=3))
    addZeemanCoupling(uc, 1, rng.normal(size=3))
    return uc


def _one_way_bond():
    # bilinear only, but a bond without its reverse: dE/dS is not the local
    # field, so the incidence gather must run
    rng = np.random.default_rng(2)
    uc = Square()
    addBasisSite(uc, [0.0, 0.0])
    M = rng.normal(size=(3, 3))
    addBilinear(uc, 1, 1, M, (1, 0))
    addBilinear(uc, 1, 1, M.T, (-1, 0))
    addBilinear(uc, 1, 1, rng.normal(size=(3, 3)), (0, 1))  # one direction only
    return uc


def _asymmetric_onsite():
    # every bond paired with its transpose, but a non-symmetric onsite matrix
    rng = np.random.default_rng(4)
    uc = Square()
    addBasisSite(uc, [0.0, 0.0])
    for d in ((1, 0), (0, 1)):
        M = rng.normal(size=(3, 3))
        addBilinear(uc, 1, 1, M, d)
        addBilinear(uc, 1, 1, M.T, (-d[0], -d[1]))
    addOnSite(uc, 1, rng.normal(size=(3, 3)))
    return uc


CASES = [
    ("random_terms", _random_cell, (3, 4)),
    ("pyro_mixed", W.pyrochlore_mixed, (2, 2, 2)),
    ("tri_j1j2", lambda: W.triangular_j1j2(1.0, 0.15), (6, 6)),
    ("tri_dm_paired", lambda: W.triangular_dm(-0.9, 0.3, [0.2, -0.1, 0.4]), (6, 6)),
    ("one_way_bond", _one_way_bond, (5, 4)),
    ("asymmetric_onsite", _asymmetric_onsite, (4, 4)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_grad_spins_vs_finite_differences(case):
    name, mk, shape = case
    uc = mk()
    lat = csmc.Lattice(shape, uc, 1.0, seed=3)
    T = orc.Tables(uc, shape)
    s = lat._spins.copy()
    b = csmc.SpinBatch([lat], precision=64)
    b.set_spins(0, s)
    g = b.grad_spins()[0]
    fd = np.zeros_like(g)
    for i in range(T.N):
        for a in range(3):
            sp, sm = s.copy(), s.copy()
            sp[i, a] += EPS
            sm[i, a] -= EPS
            fd[i, a] = (orc.total_energy(T, sp) - orc.total_energy(T, sm)) / (2 * EPS)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-7), np.max(np.abs(g - fd))


def _perturbed_energy(uc, shape, s, group, t, idx, eps):
    u = copy.deepcopy(uc)
    terms = getattr(u, group)
    term = list(terms[t])
    pos = {"field": 1, "onsite": 1, "bilinear": 2, "cubic": 3, "quartic": 4}[group]
    M = np.array(term[pos], dtype=float)
    if group in ("cubic", "quartic"):
        flat = M.reshape(-1, order="F")  # csmc_unit_cell layout: column-major
        flat[idx] += eps
        M = flat.reshape(M.shape, order="F")
    else:
        flat = M.reshape(-1)
        flat[idx] += eps
        M = flat.reshape(M.shape)
    term[pos] = M
    terms[t] = tuple(term)
    return orc.total_energy(orc.Tables(u, shape), s)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_grad_couplings_vs_finite_differences(case):
    name, mk, shape = case
    uc = mk()
    lat = csmc.Lattice(shape, uc, 1.0, seed=4)
    s = lat._spins.copy()
    b = csmc.SpinBatch([lat], precision=64)
    b.set_spins(0, s)
    g = b.grad_couplings()[0]
    off = 0
    for group, width in (("field", 3), ("onsite", 9), ("bilinear", 9), ("cubic", 27), ("quartic", 81)):
        for t in range(len(getattr(uc, group))):
            for idx in range(width):
                fd = (_perturbed_energy(uc, shape, s, group, t, idx, EPS)
                      - _perturbed_energy(uc, shape, s, group, t, idx, -EPS)) / (2 * EPS)
                assert abs(g[off] - fd) < 1e-6 * max(1.0, abs(fd)), (group, t, idx, g[off], fd)
                off += 1
    assert off == len(g)


def test_j1_j2_gradients_batch():
    # dE/dJ1 and dE/dJ2 of the triangular J1-J2 model as sums of dE/dM over
    # the J1 and J2 terms (chain rule through M = J I), for a batch of
    # configurations; equal to sum over bonds of s_i.s_j
    uc = W.triangular_j1j2(1.0, 0.1)
    lat = csmc.Lattice((12, 12), uc, 1.0, seed=5)
    n = 16
    rng = np.random.default_rng(1)
    b = csmc.SpinBatch([lat] * n, precision=64)
    confs = [random_spins(lat.size, 1.0, rng) for _ in range(n)]
    for r in range(n):
        b.set_spins(r, confs[r])
    g = b.grad_couplings()
    uc1 = W.triangular_j1j2(1.0, 0.0)
    uc1.basis.append(np.zeros(2))
    T1 = orc.Tables(uc1, (12, 12))
    for r in range(n):
        dJ1 = sum(np.trace(g[r, 9 * t:9 * t + 9].reshape(3, 3)) for t in range(6))
        dJ2 = sum(np.trace(g[r, 9 * t:9 * t + 9].reshape(3, 3)) for t in range(6, 12))
        assert abs(dJ1 - orc.total_energy(T1, confs[r])) < 1e-9 * lat.size
        s = confs[r]
        bonds2 = 0.0
        tab = lat.tables()["bsite"]
        for i in range(lat.size):
            for t in range(6, 12):
                bonds2 += 0.5 * float(np.dot(s[i], s[tab[i, t] - 1]))
        assert abs(dJ2 - bonds2) < 1e-9 * lat.size


def test_gradient_descent_reaches_deterministic_energy():
    # projected gradient descent on the sphere and deterministic updates
    # both relax to a local minimum; from the same start on the FM square
    # lattice both reach the ground state -2.1 per site
    uc = W.square_heisenberg(-1.0, (0, 0, 0.1))
    lat = csmc.Lattice((8, 8), uc, 1.0, seed=8)
    rng = np.random.default_rng(2)
    start = np.tile([0.0, 0.0, 1.0], (lat.size, 1)) + 0.4 * rng.normal(size=(lat.size, 3))
    start /= np.linalg.norm(start, axis=1, keepdims=True)
    gd = csmc.SpinBatch([lat], precision=64)
    det = csmc.SpinBatch([lat], precision=64)
    gd.set_spins(0, start)
    det.set_spins(0, start)
    E0 = gd.total_energy()[0] / lat.size
    gd.gradient_descent(0.05, 3000)
    det.sweep(csmc.DETERMINISTIC, n_sweeps=500)
    Egd = gd.total_energy()[0] / lat.size
    Edet = det.total_energy()[0] / lat.size
    assert E0 > -2.0
    assert abs(Egd + 2.1) < 1e-9 and abs(Edet + 2.1) < 1e-9, (Egd, Edet)


def test_autograd_function_matches_adjoint_and_fd():
    # csmc.autodiff.total_energy: the torch backward equals the adjoint
    # kernel and central differences of the oracle energy, with weights
    torch = pytest.importorskip("torch")
    from csmc import autodiff
    uc = _random_cell()
    shape = (3, 4)
    lat = csmc.Lattice(shape, uc, 1.0, seed=9)
    T = orc.Tables(uc, shape)
    rng = np.random.default_rng(4)
    n = 3
    confs = np.concatenate([random_spins(lat.size, 1.0, rng) for _ in range(n)])
    b = csmc.SpinBatch([lat] * n, precision=64)
    S = torch.tensor(confs, requires_grad=True)
    E = autodiff.total_energy(b, S)
    for r in range(n):
        assert abs(E[r].item() - orc.total_energy(T, confs[r * lat.size:(r + 1) * lat.size])) < 1e-10
    w = torch.tensor([0.5, -1.0, 2.0], dtype=torch.float64)
    (E * w).sum().backward()
    g = S.grad.numpy()
    for r in range(n):
        s = confs[r * lat.size:(r + 1) * lat.size]
        i, a = 5, 1
        sp, sm = s.copy(), s.copy()
        sp[i, a] += EPS
        sm[i, a] -= EPS
        fd = (orc.total_energy(T, sp) - orc.total_energy(T, sm)) / (2 * EPS)
        assert abs(g[r * lat.size + i, a] - w[r].item() * fd) < 1e-6 * max(1.0, abs(fd))


def test_fit_couplings_by_gradient():
    # dE/dJ drives a least-squares fit of J2 on the triangular J1-J2 model:
    # energies of 24 configurations at J2 = 0.13 are the targets, the fit
    # starts from J2 = 0 and must recover 0.13
    torch = pytest.importorskip("torch")
    from csmc import autodiff
    uc = W.triangular_j1j2(1.0, 0.13)
    lat = csmc.Lattice((12, 12), uc, 1.0, seed=5)
    n = 24
    rng = np.random.default_rng(6)
    confs = np.concatenate([random_spins(lat.size, 1.0, rng) for _ in range(n)])
    b = csmc.SpinBatch([lat] * n, precision=64)
    S = torch.tensor(confs)
    target = autodiff.total_energy(b, S).detach()
    I3 = torch.eye(3, dtype=torch.float64).reshape(9)
    J2 = torch.zeros(1, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([J2], lr=1.0, max_iter=50, line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        theta = torch.cat([I3.repeat(6), (J2 * I3).repeat(6)])
        loss = ((autodiff.total_energy(b, S, theta) - target) ** 2).mean()
        loss.backward()
        return loss

    opt.step(closure)
    assert abs(J2.item() - 0.13) < 1e-8, J2.item()


def test_batched_gradients_equal_single_system():
    # every system of a batch gets its own dE/dS (site offsets of the
    # concatenated output) and its own gradient step
    uc = _random_cell()
    lats = [csmc.Lattice((3, 4), uc, 1.0, seed=30 + k) for k in range(3)]
    b = csmc.SpinBatch(lats, precision=64)
    for k, lat in enumerate(lats):
        b.set_spins(k, lat._spins)
    g = b.grad_spins()
    for k, lat in enumerate(lats):
        one = csmc.SpinBatch([lat], precision=64)
        one.set_spins(0, lat._spins)
        assert np.array_equal(g[k], one.grad_spins()[0]), k
        one.gradient_descent(0.01, 3)
        if k == len(lats) - 1:
            b.gradient_descent(0.01, 3)
            for j, lj in enumerate(lats):
                ref = csmc.SpinBatch([lj], precision=64)
                ref.set_spins(0, lj._spins)
                ref.gradient_descent(0.01, 3)
                assert np.array_equal(b.get_spins(j), ref.get_spins(0)), j


@pytest.mark.parametrize("n_steps", [1, 2, 5])
def test_fused_gradient_steps_equal_gradient_then_update(n_steps):
    """csmc_gradient_descent (one fused kernel a step, two buffers) against
    the gradient of grad_spins and the projected update made on the host:
    equal to rounding for odd and even step counts, the spins back in the
    batch's own buffer (a graph captured before still sees them)"""
    uc = W.triangular_j1j2(1.0, 0.15)
    lat = csmc.Lattice((6, 6), uc, 1.0, seed=8)
    b = csmc.SpinBatch([lat], precision=64)
    b.randomize(3)
    g_or = b.graph_sweeps(csmc.OVERRELAX, n_sweeps=1)
    eta = 0.05
    s = b.get_spins(0)
    ref = csmc.SpinBatch([lat], precision=64)
    for _ in range(n_steps):
        ref.set_spins(0, s)
        g = ref.grad_spins()[0]
        u = s / np.linalg.norm(s, axis=1, keepdims=True)
        p = g - np.sum(g * u, axis=1, keepdims=True) * u
        n = s - eta * p
        s = n / np.linalg.norm(n, axis=1, keepdims=True) * 1.0
    b.gradient_descent(eta, n_steps)
    assert np.max(np.abs(b.get_spins(0) - s)) < 1e-12
    # the captured graph works on the batch's spins after the steps
    after = b.get_spins(0)
    eager = csmc.SpinBatch([lat], precision=64)
    eager.set_spins(0, after)
    eager.sweep(csmc.OVERRELAX, n_sweeps=1, plain=True)
    g_or.launch()
    assert np.array_equal(b.get_spins(0), eager.get_spins(0))
    g_or.close()
    for x in (b, ref, eager):
        x.close()
