"""Unit cells of the workloads this engine is measured and tested on, built
with the public unit-cell API (the same calls a user of the reference makes).

  square_heisenberg        square lattice, J on (+-1,0),(0,+-1), Zeeman h
                           (README.md:36-60 example; test/latticetests.jl)
  honeycomb_kitaev         honeycomb bond-dependent J, K, Gamma, Gamma' with
                           the bond convention of test/mctests.jl:14-29
  square_dm / triangular_dm  nearest-neighbour a*I + D*[z]x with the
                           transpose on reverse bonds
  pyrochlore_mixed         pyrochlore NN Heisenberg + biquadratic quartic +
                           tetrahedron-face scalar chirality + <111> onsite
  tensor_pattern_cell      honeycomb with cubic / quartic couplings of every
                           sparsity pattern the field kernels distinguish
  triangular_j1j2          isotropic J1, J2 on the triangular lattice
"""
import math

import numpy as np

from .unitcell import (Honeycomb, Pyrochlore, Square, Triangular, addBilinear, addCubic, addOnSite,
                       addQuartic, addZeemanCoupling)

I3 = np.eye(3)
ZX = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])  # [z]x


def square_heisenberg(J=-1.0, h=(0.0, 0.0, 0.1)):
    uc = Square()
    M = J * I3
    for d in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        addBilinear(uc, 1, 1, M, d)
    addZeemanCoupling(uc, 1, np.array(h, dtype=float))
    return uc


def kitaev_matrices(J=0.0, K=-1.0, G=0.0, Gp=0.0):
    Jx = np.array([[K + J, Gp, Gp], [Gp, J, G], [Gp, G, J]])
    Jy = np.array([[J, Gp, G], [Gp, K + J, Gp], [G, Gp, J]])
    Jz = np.array([[J, G, Gp], [G, J, Gp], [Gp, Gp, K + J]])
    return Jx, Jy, Jz


def honeycomb_kitaev(J=1.0, K=-2.0, G=0.0, Gp=0.0, h=None):
    uc = Honeycomb()
    Jx, Jy, Jz = kitaev_matrices(J, K, G, Gp)
    addBilinear(uc, 1, 2, Jx, (0, -1))
    addBilinear(uc, 1, 2, Jy, (1, -1))
    addBilinear(uc, 1, 2, Jz, (0, 0))
    if h is not None:
        addZeemanCoupling(uc, 1, np.array(h, dtype=float))
        addZeemanCoupling(uc, 2, np.array(h, dtype=float))
    return uc


def kgg_test_cell():
    """the K-Gamma-Gamma' cell of test/mctests.jl:2-33"""
    h = 0.1 * np.array([1.0, 1.0, 1.0]) / math.sqrt(3)
    return honeycomb_kitaev(J=0.0, K=-1.0, G=0.2, Gp=-0.02, h=h)


def _dm(a, D):
    return a * I3 + D * ZX


def square_dm(a, D, h):
    uc = Square()
    J = _dm(a, D)
    for d in ((1, 0), (0, 1)):
        addBilinear(uc, 1, 1, J, d)
        addBilinear(uc, 1, 1, J.T, (-d[0], -d[1]))
    addZeemanCoupling(uc, 1, np.array(h, dtype=float))
    return uc


def triangular_dm(a, D, h):
    uc = Triangular()
    J = _dm(a, D)
    for d in ((1, 0), (0, 1), (1, -1)):
        addBilinear(uc, 1, 1, J, d)
        addBilinear(uc, 1, 1, J.T, (-d[0], -d[1]))
    addZeemanCoupling(uc, 1, np.array(h, dtype=float))
    return uc


PYRO_BONDS = [  # (b1, b2, offset of b2's cell); the down-tetrahedron (3,4) bond is (0,-1,1)
    (1, 2, (0, 0, 0)), (1, 3, (0, 0, 0)), (1, 4, (0, 0, 0)),
    (2, 3, (0, 0, 0)), (2, 4, (0, 0, 0)), (3, 4, (0, 0, 0)),
    (1, 2, (1, 0, 0)), (1, 3, (0, 1, 0)), (1, 4, (0, 0, 1)),
    (2, 3, (-1, 1, 0)), (2, 4, (-1, 0, 1)), (3, 4, (0, -1, 1)),
]

PYRO_FACES = [  # (b1, b2, b3, o2, o3) for the 4 up and 4 down tetrahedron faces
    (1, 2, 3, (0, 0, 0), (0, 0, 0)), (1, 2, 4, (0, 0, 0), (0, 0, 0)),
    (1, 3, 4, (0, 0, 0), (0, 0, 0)), (2, 3, 4, (0, 0, 0), (0, 0, 0)),
    (1, 2, 3, (1, 0, 0), (0, 1, 0)), (1, 2, 4, (1, 0, 0), (0, 0, 1)),
    (1, 3, 4, (0, 1, 0), (0, 0, 1)), (2, 3, 4, (-1, 1, 0), (-1, 0, 1)),
]

LOCAL_Z = [np.array(v, dtype=float) / math.sqrt(3)
           for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1])]


def levi_civita():
    e = np.zeros((3, 3, 3))
    e[0, 1, 2] = e[1, 2, 0] = e[2, 0, 1] = 1.0
    e[0, 2, 1] = e[2, 1, 0] = e[1, 0, 2] = -1.0
    return e


def pyrochlore_mixed(J1=1.0, B=-0.1, chi=0.05, A=-0.2):
    uc = Pyrochlore()
    for b1, b2, o in PYRO_BONDS:
        addBilinear(uc, b1, b2, J1 * I3, o)
    biq = np.einsum("ab,cd->abcd", I3, I3) * B
    for b1, b2, o in PYRO_BONDS:
        addQuartic(uc, b1, b2, b1, b2, biq, o, (0, 0, 0), o)
    eps = levi_civita() * chi
    for b1, b2, b3, o2, o3 in PYRO_FACES:
        addCubic(uc, b1, b2, b3, eps, o2, o3)
    for b in range(4):
        z = LOCAL_Z[b]
        addOnSite(uc, b + 1, A * np.outer(z, z))
    return uc


def tensor_pattern_cell(seed=5):
    """honeycomb cell whose cubic and quartic couplings take every sparsity
    pattern the field kernels distinguish, with random weights:
    quartic (Si.Sk)(Sj.Sl)-type (role 0 sees p = r, q = t; role 1 p = t,
    q = r), (Si.Sj)(Sk.Sl)-type with anisotropic weights (p = q, r = t), a
    dense quartic on basis 1 only (so basis 2 has an all-zero padding slot);
    a cubic with non-zeros only at distinct indices, and a dense cubic on
    basis 2 only.  Bilinear Heisenberg bonds keep the spins coupled."""
    rng = np.random.default_rng(seed)
    uc = Honeycomb()
    d = np.eye(3)
    for o in ((0, 0), (1, -1), (0, -1)):
        addBilinear(uc, 1, 2, -1.0 * I3, o)
    W = rng.normal(size=(3, 3))
    addQuartic(uc, 1, 2, 1, 2, np.einsum("pr,qt,pq->pqrt", d, d, W), (0, 0), (1, 0), (1, -1))
    V = rng.normal(size=(3, 3))
    addQuartic(uc, 2, 1, 2, 1, np.einsum("pq,rt,pr->pqrt", d, d, V), (0, 1), (-1, 1), (-1, 2))
    addQuartic(uc, 1, 1, 1, 1, 0.1 * rng.normal(size=(3, 3, 3, 3)), (1, 0), (0, 1), (1, 1))
    alt = np.zeros((3, 3, 3))
    for p in range(3):
        for q in range(3):
            for r in range(3):
                if len({p, q, r}) == 3:
                    alt[p, q, r] = rng.normal()
    addCubic(uc, 1, 2, 1, alt, (0, 0), (1, 0))
    addCubic(uc, 2, 2, 2, 0.1 * rng.normal(size=(3, 3, 3)), (1, 0), (0, 1))
    return uc


J1_DIRS = ((1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1))
J2_DIRS = ((1, 1), (-1, -1), (2, -1), (-2, 1), (-1, 2), (1, -2))


def triangular_j1j2(J1=1.0, J2=0.1):
    uc = Triangular()
    for d in J1_DIRS:
        addBilinear(uc, 1, 1, J1 * I3, d)
    for d in J2_DIRS:
        addBilinear(uc, 1, 1, J2 * I3, d)
    return uc


def geometric_temperatures(Tmin, Tmax, R):
    return Tmin * (Tmax / Tmin) ** (np.arange(R) / max(R - 1, 1))


def uneven_batch(n=256, sizes=tuple(range(16, 129, 8)), seed=3):
    """independent lattices of uneven size for one batched launch: the first
    half square, the second half triangular, side L drawn uniformly from
    `sizes`; nearest-neighbour a*I + D*[z]x (a ~ N(-1, 0.2), D ~ N(0, 0.1))
    and a Zeeman field of random direction with |h| ~ U(0, 0.5), all per
    lattice.  Returns [(unit cell, (L, L))]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        L = int(rng.choice(sizes))
        a, D = rng.normal(-1.0, 0.2), rng.normal(0.0, 0.1)
        u = rng.normal(size=3)
        h = rng.uniform(0.0, 0.5) * u / np.linalg.norm(u)
        uc = square_dm(a, D, h) if k < n // 2 else triangular_dm(a, D, h)
        out.append((uc, (L, L)))
    return out


def anneal_temperatures(T0=1.0, rate=0.9, Tmin=0.05):
    """simulated_annealing!'s temperature list for schedule T0 * rate^x
    (monte_carlo.jl:166-187): T0, schedule(1), ... while T > Tmin"""
    temps, T, x = [], T0, 1
    while T > Tmin:
        temps.append(T)
        T = T0 * rate ** x
        x += 1
    return temps
