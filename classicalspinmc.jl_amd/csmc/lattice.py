"""Lattice (lattice.jl:4-345) and the Hamiltonian functions of hamiltonian.jl,
backed by libcsmc: the per-site tables are built by the library's host
builder and every energy / field evaluation runs on the GPU in FP64.
"""
import copy
import ctypes as C
import math

import numpy as np

from . import _lib
from .unitcell import addBasisSite

_BC = {"periodic": 1, "open": 0}


def random_spin_orientation(S, rng=None):
    """random_spin_orientation (lattice.jl:306-311)."""
    rng = rng if rng is not None else np.random.default_rng()
    phi = 2.0 * math.pi * rng.random()
    z = 2.0 * rng.random() - 1.0
    r = math.sqrt(1.0 - z * z)
    return (S * (r * math.cos(phi)), S * (r * math.sin(phi)), S * z)


def get_spin(spins, point):
    """get_spin (lattice.jl:293-295); spins is the (3, N) array, point 1-based."""
    return (float(spins[0, point - 1]), float(spins[1, point - 1]), float(spins[2, point - 1]))


def set_spin(spins, newspin, point):
    """set_spin! (lattice.jl:297-301)."""
    spins[0, point - 1] = newspin[0]
    spins[1, point - 1] = newspin[1]
    spins[2, point - 1] = newspin[2]


class Lattice:
    """Lattice{D,N2,N3,N4} (lattice.jl:4-24).

    ``spins`` is a (3, N) float64 array like the reference's 3xN matrix; its
    memory is the (N, 3) array handed to the library unchanged.
    The tables (bilinear_sites, ...) are exported from the library on demand
    with the reference's layout and 1-based indices.
    """

    def __init__(self, shape, uc, S=0.5, bc="periodic", initialCondition="random", seed=None):
        shape = tuple(int(v) for v in shape)
        if len(shape) != uc.D:
            raise ValueError("shape must have one entry per lattice vector")
        if len(uc.basis) == 0:  # lattice.jl:68-70 (mutates the caller's unit cell)
            addBasisSite(uc, np.zeros(uc.D))
        self.S = S
        self.bc = bc
        self.shape = shape
        self.unit_cell = uc
        self.size = int(np.prod(shape)) * len(uc.basis)
        self._periodic = _BC.get(bc, 2)
        if self._periodic == 2 and not (uc.bilinear or uc.cubic or uc.quartic):
            # the reference only calls BC() while walking terms
            self._periodic = 1
        self._info = _lib.LatticeInfo()
        self._tables = None
        self._dev = {}  # context uid -> (context, lattice id)
        self._eval = {}  # context uid -> FP64 one-system batch id
        pos = np.zeros((self.size, uc.D))
        self._build(pos=pos)
        self.site_positions = pos.T.copy()  # D x N as in lattice.jl:12
        rng = np.random.default_rng(seed)
        self._spins = np.zeros((self.size, 3))
        if initialCondition == "random" or initialCondition == ":random":
            # random_spin_orientation per site, vectorised: the same two
            # uniforms per site, in site order, as the scalar loop draws
            u = rng.random(2 * self.size).reshape(self.size, 2)
            phi = 2.0 * math.pi * u[:, 0]
            z = 2.0 * u[:, 1] - 1.0
            r = np.sqrt(1.0 - z * z)
            self._spins[:, 0] = S * (r * np.cos(phi))
            self._spins[:, 1] = S * (r * np.sin(phi))
            self._spins[:, 2] = S * z
        elif initialCondition == "fm" or initialCondition == ":fm":
            self._spins[:] = random_spin_orientation(S, rng)
        else:
            raise ValueError("initialCondition must be 'random' or 'fm'")

    # -- spins as the reference's 3xN matrix ---------------------------------
    @property
    def spins(self):
        return self._spins.T

    @spins.setter
    def spins(self, value):
        v = np.asarray(value, dtype=np.float64)
        if v.shape == (3, self.size):
            self._spins[:] = v.T
        elif v.shape == (self.size, 3):
            self._spins[:] = v
        elif v.shape == (3,):
            self._spins[:] = v
        else:
            raise ValueError("spins must be 3xN")

    def __deepcopy__(self, memo):
        new = copy.copy(self)
        new._spins = self._spins.copy()
        new._dev = self._dev  # tables are immutable: share device copies
        new._eval = {}
        return new

    # -- tables --------------------------------------------------------------
    def _build(self, **out):
        s, keep = self.unit_cell._pack()
        shape = np.ascontiguousarray(np.array(self.shape, dtype=np.int32))
        args = [out.get(k) for k in ("pos", "field", "onsite", "bsite", "bmat", "csite", "ctens",
                                     "qsite", "qtens", "colour")]
        _lib.check(_lib.lib().csmc_build_tables(
            C.byref(s), _lib.ptr(shape), self._periodic, float(self.S), C.byref(self._info),
            *[_lib.ptr(a) for a in args]))

    @property
    def info(self):
        return self._info

    def tables(self):
        """Reference-layout tables (lattice.jl:15-22), 1-based site indices."""
        if self._tables is None:
            N, D = self.size, self.unit_cell.D
            n2, n3, n4 = self._info.n2, self._info.n3, self._info.n4
            t = dict(
                pos=np.zeros((N, D)), field=np.zeros((N, 3)), onsite=np.zeros((N, 9)),
                bsite=np.zeros((N, max(n2, 0)), np.int64), bmat=np.zeros((N, n2, 9)),
                csite=np.zeros((N, n3, 2), np.int64), ctens=np.zeros((N, n3, 27)),
                qsite=np.zeros((N, n4, 3), np.int64), qtens=np.zeros((N, n4, 81)),
                colour=np.zeros(N, np.int32))
            self._build(**t)
            self._tables = t
        return self._tables

    @property
    def bilinear_sites(self):
        return self.tables()["bsite"]

    @property
    def field(self):
        return self.tables()["field"]

    @property
    def onsite(self):
        return self.tables()["onsite"]

    def colouring(self):
        return self.tables()["colour"]

    # -- device residency -------------------------------------------------------
    def device_id(self, ctx=None):
        ctx = ctx or _lib.default_context()
        key = ctx.uid
        if key not in self._dev or not self._dev[key][0].h:
            s, keep = self.unit_cell._pack()
            shape = np.ascontiguousarray(np.array(self.shape, dtype=np.int32))
            lid = C.c_int32()
            _lib.check(_lib.lib().csmc_lattice_create(ctx.h, C.byref(s), _lib.ptr(shape),
                                                      self._periodic, float(self.S), C.byref(lid)))
            self._dev[key] = (ctx, lid.value)
        return self._dev[key][1]

    def update_couplings(self, uc):
        """replace the couplings of this lattice by those of `uc`, a unit
        cell with the same bases, offsets and term order (for fitting
        couplings without rebuilding tables); device copies are updated in
        place, so batches holding this lattice see the new values"""
        s, keep = uc._pack()
        for ctx, lid in self._dev.values():
            if ctx.h:
                _lib.check(_lib.lib().csmc_lattice_update_couplings(ctx.h, lid, C.byref(s)))
        self.unit_cell = uc
        self._tables = None

    def _eval_batch(self, ctx=None, upload=True):
        ctx = ctx or _lib.default_context()
        key = ctx.uid
        if key not in self._eval:
            lid = np.array([self.device_id(ctx)], dtype=np.int32)
            bid = C.c_int32()
            _lib.check(_lib.lib().csmc_batch_create(ctx.h, 1, _lib.ptr(lid), _lib.FP64, C.byref(bid)))
            self._eval[key] = bid.value
        bid = self._eval[key]
        if upload:
            _lib.check(_lib.lib().csmc_set_spins(ctx.h, bid, 0, _lib.ptr(self._spins)))
        return ctx, bid

    def _site_and_partners(self, i):
        """0-based site i and every site a term of it names (the reference
        tables: bilinear, cubic and quartic partners; 0 = none)"""
        t = self.tables()
        parts = [np.array([i + 1], dtype=np.int64)]
        for k in ("bsite", "csite", "qsite"):
            a = np.asarray(t[k][i], dtype=np.int64).reshape(-1)
            parts.append(a[a > 0])
        return np.unique(np.concatenate(parts)) - 1

    def _eval_site(self, i):
        """local field and site energy of 0-based site i from the current host
        spins: only the site and its partners are written to the device, so
        a call costs the same at any lattice size (csmc_site_eval)"""
        if not 0 <= i < self.size:
            raise IndexError(f"site {i + 1} out of 1..{self.size}")
        ctx, bid = self._eval_batch(upload=False)
        sites = self._site_and_partners(i)
        vals = np.ascontiguousarray(self._spins[sites], dtype=np.float64)
        at = np.array([i], dtype=np.int64)
        H = np.zeros(3)
        e = np.zeros(1)
        _lib.check(_lib.lib().csmc_site_eval(ctx.h, bid, 0, len(sites), _lib.ptr(sites), _lib.ptr(vals), 1,
                                             _lib.ptr(at), _lib.ptr(H), _lib.ptr(e)))
        return H, float(e[0])


def get_local_field(lattice, point):
    """get_local_field (hamiltonian.jl:3-67); point is 1-based."""
    H, _ = lattice._eval_site(point - 1)
    return tuple(float(v) for v in H)


def local_fields(lattice):
    """get_local_field at every site, (N, 3)."""
    ctx, bid = lattice._eval_batch()
    H = np.zeros((lattice.size, 3))
    _lib.check(_lib.lib().csmc_local_field(ctx.h, bid, 0, _lib.ptr(H)))
    return H


def site_energies(lattice):
    """energy(lattice, point) (hamiltonian.jl:139-196) at every site."""
    ctx, bid = lattice._eval_batch()
    e = np.zeros(lattice.size)
    _lib.check(_lib.lib().csmc_site_energy(ctx.h, bid, 0, _lib.ptr(e)))
    return e


def energy(lattice, point):
    """energy(lattice, point) (hamiltonian.jl:139-196); point 1-based."""
    return lattice._eval_site(point - 1)[1]


def total_energy(lattice, parts=False):
    """total_energy (hamiltonian.jl:70-132), FP64 on the GPU."""
    ctx, bid = lattice._eval_batch()
    E = np.zeros(1)
    P = np.zeros(5)
    _lib.check(_lib.lib().csmc_total_energy(ctx.h, bid, _lib.ptr(E), _lib.ptr(P)))
    return (float(E[0]), P) if parts else float(E[0])


def energy_density(lattice):
    """energy_density (hamiltonian.jl:134-136)."""
    return total_energy(lattice) / lattice.size


def get_magnetization(lattice):
    """get_magnetization (observables.jl:12-18): |sum_i s_i|, FP64 on the GPU."""
    ctx, bid = lattice._eval_batch()
    M = np.zeros(3)
    _lib.check(_lib.lib().csmc_magnetization(ctx.h, bid, _lib.ptr(M)))
    return float(np.linalg.norm(M))


def compute_equal_time_correlations(lattice, ks):
    """compute_equal_time_correlations (spin_correlations.jl:6-43): the 9 x n_k
    matrix Suv (rows xx xy xz yx yy yz zx zy zz) of equal-time spin
    correlations at the momenta ks (D x n_k), divided by lattice.size.  FP64
    on the GPU (phases and two GEMMs)."""
    ks = np.asarray(ks, dtype=np.float64)
    if ks.ndim != 2 or ks.shape[0] != lattice.unit_cell.D:
        raise ValueError("ks must be a D x n_k matrix")
    ctx, bid = lattice._eval_batch()
    kt = np.ascontiguousarray(ks.T)
    out = np.zeros((ks.shape[1], 9))
    _lib.check(_lib.lib().csmc_spin_correlations(ctx.h, bid, _lib.ptr(kt), ks.shape[1], _lib.ptr(out)))
    return out.T.copy()
