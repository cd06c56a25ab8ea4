"""Model description: UnitCell and its term builders (unit_cell.jl:3-75),
InteractionMatrix helpers (interaction_matrix.jl) and the preset unit cells
(bravais.jl:1-51).

Python names drop Julia's `!`: addBilinear!(uc, ...) is addBilinear(uc, ...).
Basis indices are 1-based as in the reference.  Tensors are numpy arrays
indexed like the Julia ones (M[a, b, c] == M[a+1, b+1, c+1] in Julia).
"""
import ctypes as C
import math

import numpy as np

from . import _lib


def InteractionMatrix(J):
    """InteractionMatrix(J) (interaction_matrix.jl:14-18): a 3x3 float64
    matrix (the nine fields m11..m33 in row-major order)."""
    M = np.array(J, dtype=np.float64).reshape(3, 3)
    return M


def IMToMatrix(J):
    """IMToMatrix (interaction_matrix.jl:21-25)."""
    return np.array(J, dtype=np.float64).reshape(3, 3)


def transposeJ(J):
    """transposeJ (interaction_matrix.jl:27-31)."""
    return np.array(J, dtype=np.float64).reshape(3, 3).T.copy()


class UnitCell:
    """UnitCell{D} (unit_cell.jl:3-18)."""

    def __init__(self, *lattice_vectors):
        self.lattice_vectors = tuple(np.array(a, dtype=np.float64) for a in lattice_vectors)
        D = len(self.lattice_vectors)
        if not 1 <= D <= 3:
            raise ValueError("UnitCell needs 1 to 3 lattice vectors")
        for a in self.lattice_vectors:
            if a.shape != (D,):
                raise ValueError("lattice vectors must have D components")
        self.basis = []
        self.field = []      # (b, h)
        self.onsite = []     # (b, M)
        self.bilinear = []   # (b1, b2, M, offset)
        self.cubic = []      # (b1, b2, b3, M, o2, o3)
        self.quartic = []    # (b1, b2, b3, b4, M, o2, o3, o4)

    @property
    def D(self):
        return len(self.lattice_vectors)

    def _off(self, o):
        o = tuple(int(v) for v in (o if o is not None else (0,) * self.D))
        if len(o) != self.D:
            raise ValueError(f"offset must have {self.D} components")
        return o

    # -- flattened C view ---------------------------------------------------
    def _pack(self):
        """csmc_unit_cell for this unit cell; returns (struct, keepalive)."""
        D = self.D
        keep = []

        def f64(a, shape):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape)) if np.size(a) else np.zeros(1)
            keep.append(arr)
            return arr.ctypes.data_as(C.POINTER(C.c_double))

        def i32(a, shape):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(shape)) if np.size(a) else np.zeros(1, np.int32)
            keep.append(arr)
            return arr.ctypes.data_as(C.POINTER(C.c_int32))

        s = _lib.UnitCellStruct()
        s.dim = D
        s.n_basis = len(self.basis)
        s.lattice_vectors = f64(np.stack(self.lattice_vectors), (D, D))
        s.basis = f64(np.array(self.basis).reshape(-1), (-1,))
        s.n_field = len(self.field)
        s.field_b = i32([b for b, _ in self.field], (-1,))
        s.field_h = f64([h for _, h in self.field], (-1,))
        s.n_onsite = len(self.onsite)
        s.onsite_b = i32([b for b, _ in self.onsite], (-1,))
        s.onsite_m = f64([np.asarray(M).reshape(9) for _, M in self.onsite], (-1,))
        s.n_bilinear = len(self.bilinear)
        s.bilinear_b = i32([(b1, b2) for b1, b2, _, _ in self.bilinear], (-1,))
        s.bilinear_m = f64([np.asarray(M).reshape(9) for _, _, M, _ in self.bilinear], (-1,))
        s.bilinear_off = i32([o for _, _, _, o in self.bilinear], (-1,))
        s.n_cubic = len(self.cubic)
        s.cubic_b = i32([t[:3] for t in self.cubic], (-1,))
        s.cubic_t = f64([np.asarray(t[3]).flatten(order="F") for t in self.cubic], (-1,))
        s.cubic_off = i32([(t[4], t[5]) for t in self.cubic], (-1,))
        s.n_quartic = len(self.quartic)
        s.quartic_b = i32([t[:4] for t in self.quartic], (-1,))
        s.quartic_t = f64([np.asarray(t[4]).flatten(order="F") for t in self.quartic], (-1,))
        s.quartic_off = i32([(t[5], t[6], t[7]) for t in self.quartic], (-1,))
        return s, keep

    def n_params(self):
        """length of the coupling gradient vector (field, onsite, bilinear,
        cubic, quartic in call order)."""
        return (3 * len(self.field) + 9 * len(self.onsite) + 9 * len(self.bilinear)
                + 27 * len(self.cubic) + 81 * len(self.quartic))

    def params(self):
        """every coupling as one float64 vector, in the order and layout of
        the coupling gradient (cubic / quartic tensors column-major)"""
        parts = [np.asarray(h, float).reshape(3) for _, h in self.field]
        parts += [np.asarray(M, float).reshape(9) for _, M in self.onsite]
        parts += [np.asarray(t[2], float).reshape(9) for t in self.bilinear]
        parts += [np.asarray(t[3], float).flatten(order="F") for t in self.cubic]
        parts += [np.asarray(t[4], float).flatten(order="F") for t in self.quartic]
        return np.concatenate(parts) if parts else np.zeros(0)

    def with_params(self, theta):
        """a copy of this unit cell with the couplings replaced by `theta`
        (layout of params()); bases, offsets and term order are kept"""
        theta = np.asarray(theta, dtype=np.float64).reshape(-1)
        if theta.size != self.n_params():
            raise ValueError(f"expected {self.n_params()} couplings, got {theta.size}")
        u = UnitCell(*self.lattice_vectors)
        u.basis = [np.array(b) for b in self.basis]
        k = 0

        def take(n, shape, order="C"):
            nonlocal k
            v = theta[k:k + n].reshape(shape, order=order).copy()
            k += n
            return v

        u.field = [(b, take(3, (3,))) for b, _ in self.field]
        u.onsite = [(b, take(9, (3, 3))) for b, _ in self.onsite]
        u.bilinear = [(t[0], t[1], take(9, (3, 3)), t[3]) for t in self.bilinear]
        u.cubic = [t[:3] + (take(27, (3, 3, 3), "F"),) + t[4:] for t in self.cubic]
        u.quartic = [t[:4] + (take(81, (3, 3, 3, 3), "F"),) + t[5:] for t in self.quartic]
        return u


def addBasisSite(uc, site):
    """addBasisSite! (unit_cell.jl:23-25)."""
    site = np.array(site, dtype=np.float64)
    if site.shape != (uc.D,):
        raise ValueError(f"basis site must have {uc.D} components")
    uc.basis.append(site)


def addZeemanCoupling(uc, b1, h):
    """addZeemanCoupling! (unit_cell.jl:30-32); zero fields are kept."""
    uc.field.append((int(b1), np.array(h, dtype=np.float64).reshape(3)))


def addOnSite(uc, b1, M):
    """addOnSite! (unit_cell.jl:37-41); an all-zero matrix is dropped."""
    M = InteractionMatrix(M)
    if np.any(M != 0):
        uc.onsite.append((int(b1), M))


def addBilinear(uc, b1, b2, M, offset=None):
    """addBilinear! (unit_cell.jl:46-52); an all-zero matrix is dropped."""
    M = InteractionMatrix(M)
    if np.any(M != 0):
        uc.bilinear.append((int(b1), int(b2), M, uc._off(offset)))


def addCubic(uc, b1, b2, b3, M, o2=None, o3=None):
    """addCubic! (unit_cell.jl:57-63); an all-zero tensor is dropped."""
    M = np.array(M, dtype=np.float64).reshape(3, 3, 3)
    if np.any(M != 0):
        uc.cubic.append((int(b1), int(b2), int(b3), M, uc._off(o2), uc._off(o3)))


def addQuartic(uc, b1, b2, b3, b4, M, o2=None, o3=None, o4=None):
    """addQuartic! (unit_cell.jl:68-75); an all-zero tensor is dropped."""
    M = np.array(M, dtype=np.float64).reshape(3, 3, 3, 3)
    if np.any(M != 0):
        uc.quartic.append((int(b1), int(b2), int(b3), int(b4), M, uc._off(o2), uc._off(o3),
                           uc._off(o4)))


# -- presets (bravais.jl:1-51); they carry no Hamiltonian terms ------------

def Triangular():
    return UnitCell([1.0, 0.0], [math.cos(math.pi / 3), math.sin(math.pi / 3)])


def Square():
    return UnitCell([1.0, 0.0], [0.0, 1.0])


def FCC():
    return UnitCell([0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0])


def _pyro_basis(a=1.0):
    return [a * np.array(v, dtype=np.float64) / 8
            for v in ([1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1])]


def Pyrochlore():
    uc = FCC()
    for site in _pyro_basis():
        addBasisSite(uc, site)
    return uc


def BreathingPyrochlore(a=1.01):
    uc = FCC()
    for site in _pyro_basis(a):
        addBasisSite(uc, site)
    return uc


def Honeycomb():
    uc = Triangular()
    addBasisSite(uc, [0.0, 0.0])
    addBasisSite(uc, [0.0, 1.0 / math.sqrt(3)])
    return uc
