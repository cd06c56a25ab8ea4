"""The lattices of the golden fixtures (tests/golden/golden.npz): small
cells that exercise every term kind, both boundary conditions and the
reference's own test cells."""
import numpy as np

from csmc import workloads as W
from csmc.unitcell import Honeycomb, addBilinear, addCubic, addOnSite, addQuartic, addZeemanCoupling


def random_terms_cell():
    rng = np.random.default_rng(21)
    uc = Honeycomb()
    addCubic(uc, 2, 1, 2, rng.normal(size=(3, 3, 3)), (1, 0), (0, -1))
    addQuartic(uc, 1, 2, 2, 1, rng.normal(size=(3, 3, 3, 3)), (0, 0), (1, -1), (-1, 1))
    addBilinear(uc, 2, 1, rng.normal(size=(3, 3)), (1, 1))
    addBilinear(uc, 2, 2, rng.normal(size=(3, 3)), (0, 2))
    addOnSite(uc, 1, rng.normal(size=(3, 3)))
    addOnSite(uc, 2, rng.normal(size=(3, 3)))
    addZeemanCoupling(uc, 2, rng.normal(size=3))
    addZeemanCoupling(uc, 1, rng.normal(size=3))
    return uc


# (name, unit cell factory, shape, S, bc)
CASES = [
    ("square_fm", lambda: W.square_heisenberg(-1.0, (0, 0, 0.1)), (4, 4), 1.0, "periodic"),
    ("honeycomb_kitaev", lambda: W.honeycomb_kitaev(1.0, -2.0), (4, 4), 0.5, "periodic"),
    ("mctests_kgg", W.kgg_test_cell, (4, 4), 1.0, "periodic"),
    ("kgg_open", W.kgg_test_cell, (5, 3), 1.0, "open"),
    ("triangular_dm", lambda: W.triangular_dm(-0.9, -0.05, [0.3, 0.1, 0.0]), (6, 6), 1.0, "periodic"),
    ("pyrochlore_mixed", W.pyrochlore_mixed, (2, 2, 2), 1.0, "periodic"),
    ("triangular_j1j2", lambda: W.triangular_j1j2(1.0, 0.13), (6, 6), 1.0, "periodic"),
    ("random_terms", random_terms_cell, (3, 4), 1.0, "periodic"),
]


def make_cell(mk):
    """the unit cell of a case with the default basis site the Lattice
    constructor adds when there is none (lattice.jl:68-70)"""
    uc = mk()
    if len(uc.basis) == 0:
        uc.basis.append(np.zeros(uc.D))
    return uc


def colour_order(lat):
    """1-based sites in sweep order: colour classes in turn, each in
    reference order (the order the device and the oracle replay)"""
    col = lat.colouring()
    order = []
    for c in range(int(col.max()) + 1):
        order.extend((np.nonzero(col == c)[0] + 1).tolist())
    return np.array(order, dtype=np.int64)
