"""ctypes wrapper of the CPU FP64 oracle (csmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker, never by the product.

The unit cell argument is any object with the attributes of the reference's
UnitCell (lattice_vectors, basis, field, onsite, bilinear, cubic, quartic;
unit_cell.jl:3-18) with 1-based basis indices.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_long)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_bench_build():
    """Load liboracle_bench.so (the same source at -O3 for x86-64-v3) instead
    of the checker build; for bench.py's CPU baseline only, before the first
    oracle call of the process."""
    global _SO
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _SO = os.path.join(_HERE, "liboracle_bench.so")
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = C.CDLL(_SO)
        _lib.orc_total_energy.restype = C.c_double
        _lib.orc_site_energy.restype = C.c_double
        _lib.orc_metropolis_coloured.restype = C.c_long
        _lib.orc_metropolis_random.restype = C.c_long
        _lib.orc_adapt_sigma.restype = C.c_double
        _lib.orc_adapt_sigma.argtypes = [C.c_double, C.c_double, C.c_double]
        _lib.orc_magnetization.restype = C.c_double
        _lib.orc_binner_var.restype = C.c_double
        _lib.orc_binner_std_error.restype = C.c_double
        _lib.orc_binner_size.restype = C.c_size_t
        _lib.orc_pt_reference_run.restype = C.c_long
        _lib.orc_metropolis_constraint_coloured.restype = C.c_long
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class OrcLat(C.Structure):
    _fields_ = [("N", C.c_long), ("n2", C.c_int), ("n3", C.c_int), ("n4", C.c_int),
                ("bsite", _L), ("bmat", _D), ("csite", _L), ("ctens", _D),
                ("qsite", _L), ("qtens", _D), ("field", _D), ("onsite", _D)]


class Tables:
    """The reference Lattice tables (lattice.jl:10-22) built by the oracle."""

    def __init__(self, uc, shape, periodic=True, search=0):
        D = len(uc.lattice_vectors)
        nb = len(uc.basis)
        shape = np.array(shape, dtype=np.intc)
        N = int(np.prod(shape)) * nb
        n2, n3, n4 = len(uc.bilinear), len(uc.cubic), len(uc.quartic)
        self.N, self.n2, self.n3, self.n4, self.dim = N, n2, n3, n4, D
        self.pos = np.zeros((N, D))
        self.field = np.zeros((N, 3))
        self.onsite = np.zeros((N, 9))
        self.bsite = np.zeros((N, n2), dtype=np.int_)
        self.bmat = np.zeros((N, n2, 9))
        self.csite = np.zeros((N, n3, 2), dtype=np.int_)
        self.ctens = np.zeros((N, n3, 27))
        self.qsite = np.zeros((N, n4, 3), dtype=np.int_)
        self.qtens = np.zeros((N, n4, 81))

        def f(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
            return a if a.size else np.zeros(1)

        def i(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.intc).reshape(-1))
            return a if a.size else np.zeros(1, np.intc)

        keep = dict(
            avec=f(np.stack([np.asarray(a, float) for a in uc.lattice_vectors])),
            basis=f(np.array(uc.basis, dtype=float)),
            fb=i([t[0] for t in uc.field]), fh=f([np.asarray(t[1]) for t in uc.field]),
            ob=i([t[0] for t in uc.onsite]), om=f([np.asarray(t[1]).reshape(9) for t in uc.onsite]),
            b2=i([t[:2] for t in uc.bilinear]), m2=f([np.asarray(t[2]).reshape(9) for t in uc.bilinear]),
            o2=i([t[3] for t in uc.bilinear]),
            b3=i([t[:3] for t in uc.cubic]),
            t3=f([np.asarray(t[3]).flatten(order="F") for t in uc.cubic]),
            o3=i([(t[4], t[5]) for t in uc.cubic]),
            b4=i([t[:4] for t in uc.quartic]),
            t4=f([np.asarray(t[4]).flatten(order="F") for t in uc.quartic]),
            o4=i([(t[5], t[6], t[7]) for t in uc.quartic]))
        k = keep
        rc = lib().orc_build_tables(
            D, _p(shape, _I), 1 if periodic else 0, search, _p(k["avec"], _D), nb, _p(k["basis"], _D),
            len(uc.field), _p(k["fb"], _I), _p(k["fh"], _D),
            len(uc.onsite), _p(k["ob"], _I), _p(k["om"], _D),
            n2, _p(k["b2"], _I), _p(k["m2"], _D), _p(k["o2"], _I),
            n3, _p(k["b3"], _I), _p(k["t3"], _D), _p(k["o3"], _I),
            n4, _p(k["b4"], _I), _p(k["t4"], _D), _p(k["o4"], _I),
            _p(self.pos, _D), _p(self.field, _D), _p(self.onsite, _D),
            _p(self.bsite, _L), _p(self.bmat, _D), _p(self.csite, _L), _p(self.ctens, _D),
            _p(self.qsite, _L), _p(self.qtens, _D))
        self.rc = rc
        if rc != 0:
            raise OracleError(rc)
        self.S = None
        self._s = OrcLat(N, n2, n3, n4, _p(self.bsite, _L), _p(self.bmat, _D), _p(self.csite, _L),
                         _p(self.ctens, _D), _p(self.qsite, _L), _p(self.qtens, _D),
                         _p(self.field, _D), _p(self.onsite, _D))

    @property
    def ref(self):
        return C.byref(self._s)


class OracleError(RuntimeError):
    CODES = {1: "BoundsError", 2: "undef", 3: "Invalid boundary condition option", 4: "bad argument"}

    def __init__(self, rc):
        super().__init__(self.CODES.get(rc, str(rc)))
        self.rc = rc


def _spins(s):
    s = np.ascontiguousarray(np.asarray(s, dtype=np.float64))
    assert s.ndim == 2 and s.shape[1] == 3, "spins must be (N, 3)"
    return s


def local_field(T, spins, point):
    H = np.zeros(3)
    s = _spins(spins)
    lib().orc_local_field(T.ref, _p(s, _D), C.c_long(point), _p(H, _D))
    return H


def local_fields(T, spins):
    """get_local_field of every site, (N, 3), in one C loop (no GIL held)"""
    s = _spins(spins)
    H = np.zeros((T.N, 3))
    lib().orc_local_fields(T.ref, _p(s, _D), _p(H, _D))
    return H


def site_energy(T, spins, point):
    s = _spins(spins)
    return lib().orc_site_energy(T.ref, _p(s, _D), C.c_long(point))


def site_energies(T, spins):
    return np.array([site_energy(T, spins, p) for p in range(1, T.N + 1)])


def total_energy(T, spins, parts=False):
    s = _spins(spins)
    P = np.zeros(5)
    E = lib().orc_total_energy(T.ref, _p(s, _D), _p(P, _D))
    return (E, P) if parts else E


def overrelax(T, spins, order):
    """in place; order = 1-based sites"""
    o = np.ascontiguousarray(np.asarray(order, dtype=np.int_))
    lib().orc_overrelax_order(T.ref, _p(spins, _D), _p(o, _L), C.c_long(len(o)))


def deterministic(T, spins, order, S):
    o = np.ascontiguousarray(np.asarray(order, dtype=np.int_))
    lib().orc_deterministic_order(T.ref, _p(spins, _D), _p(o, _L), C.c_long(len(o)), C.c_double(S))


def random_spin(S, u1, u2):
    o = np.zeros(3)
    lib().orc_random_spin(C.c_double(S), C.c_double(u1), C.c_double(u2), _p(o, _D))
    return o


def cone_move(S, spin, sigma, u1, u2):
    o = np.zeros(3)
    s = np.ascontiguousarray(np.asarray(spin, dtype=np.float64))
    lib().orc_cone_move(C.c_double(S), _p(s, _D), C.c_double(sigma), C.c_double(u1), C.c_double(u2),
                        _p(o, _D))
    return o


def philox(ctr, key):
    c = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32))
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
    o = np.zeros(4, dtype=np.uint32)
    P = C.POINTER(C.c_uint32)
    lib().orc_philox4x32_10(_p(c, P), _p(k, P), _p(o, P))
    return o


def metropolis_coloured(T, spins, order, S, temp, kind, sigma, seed, stream, sweep_ctr):
    """returns (accepted, decisions[len(order)], dE[len(order)])"""
    o = np.ascontiguousarray(np.asarray(order, dtype=np.int_))
    dec = np.zeros(len(o), dtype=np.uint8)
    dE = np.zeros(len(o))
    a = lib().orc_metropolis_coloured(
        T.ref, _p(spins, _D), _p(o, _L), C.c_long(len(o)), C.c_double(S), C.c_double(temp),
        C.c_int(kind), C.c_double(sigma), C.c_uint64(seed), C.c_uint32(stream), C.c_uint32(sweep_ctr),
        _p(dec, C.POINTER(C.c_ubyte)), _p(dE, _D))
    return a, dec, dE


def metropolis_constraint_coloured(T, spins, order, S, temp, sigma, sigma0, seed, stream, sweep_ctr, ncell,
                                   a, Q, c0, weight, lam):
    """metropolis_constraint_adaptive! over `order` (orc_metropolis_constraint_coloured):
    returns (accepted, decisions, sigma after the width law, lambda)"""
    o = np.ascontiguousarray(np.asarray(order, dtype=np.int_))
    dec = np.zeros(len(o), dtype=np.uint8)
    av = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))
    qv = np.ascontiguousarray(np.asarray(Q, dtype=np.float64).reshape(-1, 9))
    sg = C.c_double(sigma)
    lm = C.c_double(lam)
    acc = lib().orc_metropolis_constraint_coloured(
        T.ref, _p(spins, _D), _p(o, _L), C.c_long(len(o)), C.c_double(S), C.c_double(temp), C.byref(sg),
        C.c_double(sigma0), C.c_uint64(seed), C.c_uint32(stream), C.c_uint32(sweep_ctr), C.c_long(ncell),
        _p(av, _D), _p(qv, _D), C.c_int(len(av)), C.c_double(c0), C.c_double(weight), C.byref(lm),
        _p(dec, C.POINTER(C.c_ubyte)))
    return acc, dec, sg.value, lm.value


def metropolis_random(T, spins, S, temp, kind, sigma, seed, stream, sweep_ctr):
    return lib().orc_metropolis_random(T.ref, _p(spins, _D), C.c_double(S), C.c_double(temp),
                                       C.c_int(kind), C.c_double(sigma), C.c_uint64(seed),
                                       C.c_uint32(stream), C.c_uint32(sweep_ctr))


def adapt_sigma(sigma, accepted, attempts):
    return lib().orc_adapt_sigma(sigma, accepted, attempts)


def magnetization(spins):
    s = _spins(spins)
    m = np.zeros(3)
    return lib().orc_magnetization(_p(s, _D), C.c_long(len(s)), _p(m, _D)), m


def pt_exchange(temps, energy, perm, seed, xc, accepted=None):
    t = np.ascontiguousarray(np.asarray(temps, dtype=np.float64))
    assert energy.dtype == np.float64 and perm.dtype == np.intc
    acc = accepted if accepted is not None else np.zeros(len(t), dtype=np.int_)
    n = lib().orc_pt_exchange(C.c_int(len(t)), _p(t, _D), _p(energy, _D), _p(perm, _I), C.c_uint64(seed),
                              C.c_uint32(xc), _p(acc, _L))
    return n


class Binner:
    """log binner with error propagation (observables.jl:5-63 semantics)"""

    def __init__(self):
        self.buf = (C.c_char * lib().orc_binner_size())()
        lib().orc_binner_init(self.buf, 2)

    def push(self, v):
        a = np.array([v, v * v])
        lib().orc_binner_push(self.buf, _p(a, _D))

    def push2(self, a, b):
        x = np.array([a, b], dtype=np.float64)
        lib().orc_binner_push(self.buf, _p(x, _D))

    def var(self, g, lvl):
        g = np.ascontiguousarray(np.asarray(g, dtype=np.float64))
        return lib().orc_binner_var(self.buf, _p(g, _D), C.c_int(lvl))

    def std_error(self, g, lvl=None):
        g = np.ascontiguousarray(np.asarray(g, dtype=np.float64))
        if lvl is None:
            lvl = lib().orc_binner_reliable_level(self.buf)
        return lib().orc_binner_std_error(self.buf, _p(g, _D), C.c_int(lvl))

    def reliable_level(self):
        return lib().orc_binner_reliable_level(self.buf)

    def means(self):
        m = np.zeros(2)
        lib().orc_binner_means(self.buf, _p(m, _D))
        return m


def observables(bE, bM, T, N):
    out = np.zeros(8)
    lib().orc_observables(bE.buf, bM.buf, C.c_double(T), C.c_long(N), _p(out, _D))
    return out


def pt_reference_run(T, spins, temps, S, sweep0, n_sweeps, or_rate, swap_rate, seed, E, perm,
                     acc_local=None, acc_exch=None):
    """spins: (R, N, 3) C-contiguous, modified in place; returns updates"""
    t = np.ascontiguousarray(np.asarray(temps, dtype=np.float64))
    R = len(t)
    al = acc_local if acc_local is not None else np.zeros(R, dtype=np.int_)
    ax = acc_exch if acc_exch is not None else np.zeros(R, dtype=np.int_)
    return lib().orc_pt_reference_run(
        T.ref, _p(spins, _D), C.c_int(R), _p(t, _D), C.c_double(S), C.c_long(sweep0),
        C.c_long(n_sweeps), C.c_long(or_rate), C.c_long(swap_rate), C.c_uint64(seed), _p(E, _D),
        _p(perm, _I), _p(al, _L), _p(ax, _L))


def equal_time_correlations(pos, spins, ks):
    """compute_equal_time_correlations (spin_correlations.jl:6-43).
    pos: D x N site positions, spins: (N, 3), ks: D x n_k; returns the
    9 x n_k Suv matrix (rows xx xy xz yx yy yz zx zy zz), divided by N."""
    P = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).T)
    K = np.ascontiguousarray(np.asarray(ks, dtype=np.float64).T)
    s = _spins(spins)
    N, D = P.shape
    nk = K.shape[0]
    out = np.zeros((nk, 9))
    lib().orc_equal_time_correlations(_p(P, _D), C.c_long(N), C.c_int(D), _p(s, _D), _p(K, _D),
                                      C.c_long(nk), _p(out, _D))
    return out.T.copy()
