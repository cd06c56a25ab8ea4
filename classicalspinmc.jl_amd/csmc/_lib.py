"""ctypes binding of libcsmc (include/csmc.h).

The shared library is built in-tree (``make`` in ``classicalspinmc.jl_amd``)
and loaded from this directory.  There is no fallback: if the library is
missing the import raises, and every device entry point fails with the
library's own error when no HIP device is present.
"""
import atexit
import ctypes as C
import itertools
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsmc.so")

OK, EINVAL, EBOUNDS, EUNDEF, EBC, EHIP, ENOMEM, ESTATE = range(8)
FP32, FP64 = 32, 64
OVERRELAX, METROPOLIS, CONE, CONE_ADAPTIVE, DETERMINISTIC, CONSTRAINT_ADAPTIVE = range(6)


class CsmcError(RuntimeError):
    """Failure reported by libcsmc (status code + message)."""

    def __init__(self, code, msg):
        super().__init__(f"libcsmc error {code}: {msg}")
        self.code = code


class BoundsError(CsmcError, IndexError):
    """The reference raises BoundsError for this input (lattice.jl:132,138)."""


class UndefError(CsmcError, ValueError):
    """The reference would read an undef entry for this input."""


class UnitCellStruct(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("n_basis", C.c_int32),
        ("lattice_vectors", C.POINTER(C.c_double)), ("basis", C.POINTER(C.c_double)),
        ("n_field", C.c_int32), ("field_b", C.POINTER(C.c_int32)), ("field_h", C.POINTER(C.c_double)),
        ("n_onsite", C.c_int32), ("onsite_b", C.POINTER(C.c_int32)), ("onsite_m", C.POINTER(C.c_double)),
        ("n_bilinear", C.c_int32), ("bilinear_b", C.POINTER(C.c_int32)),
        ("bilinear_m", C.POINTER(C.c_double)), ("bilinear_off", C.POINTER(C.c_int32)),
        ("n_cubic", C.c_int32), ("cubic_b", C.POINTER(C.c_int32)),
        ("cubic_t", C.POINTER(C.c_double)), ("cubic_off", C.POINTER(C.c_int32)),
        ("n_quartic", C.c_int32), ("quartic_b", C.POINTER(C.c_int32)),
        ("quartic_t", C.POINTER(C.c_double)), ("quartic_off", C.POINTER(C.c_int32)),
    ]


class LatticeInfo(C.Structure):
    _fields_ = [
        ("n_sites", C.c_int64), ("dim", C.c_int32), ("n_basis", C.c_int32),
        ("n2", C.c_int32), ("n3", C.c_int32), ("n4", C.c_int32),
        ("k2", C.c_int32), ("k3", C.c_int32), ("k4", C.c_int32),
        ("n_colours", C.c_int32), ("has_self", C.c_int32), ("S", C.c_double),
    ]


class PTParams(C.Structure):
    _fields_ = [
        ("t_thermalization", C.c_int64), ("t_measurement", C.c_int64), ("probe_rate", C.c_int64),
        ("swap_rate", C.c_int64), ("overrelaxation_rate", C.c_int64),
        ("report_interval", C.c_int64), ("checkpoint_rate", C.c_int64),
        ("alg", C.c_int32), ("measure", C.c_int32),
    ]


_P = C.c_void_p
_I32, _I64, _U64, _D = C.c_int32, C.c_int64, C.c_uint64, C.c_double

# name -> argument types (every entry point returns int32 status, except
# csmc_last_error)
SIGNATURES = {
    "csmc_version": [],
    "csmc_device_count": [_P],
    "csmc_ctx_create": [_I32, _P],
    "csmc_ctx_destroy": [_P],
    "csmc_ctx_sync": [_P],
    "csmc_timing": [_P, _I32],
    "csmc_timing_read": [_P, _P, _P],
    "csmc_build_tables": [_P, _P, _I32, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "csmc_split_plan": [_P, _P, _I32, _I32, _I32, _P],
    "csmc_tile_plan": [_P, _P, _I32, _I32, _I64, _I32, _P, _P, _I32, _P],
    "csmc_batch_tiles": [_P, _I32, _P, _P],
    "csmc_site_eval": [_P, _I32, _I32, _I32, _P, _P, _I32, _P, _P, _P],
    "csmc_lattice_create": [_P, _P, _P, _I32, _D, _P],
    "csmc_lattice_info_get": [_P, _I32, _P],
    "csmc_lattice_tables": [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "csmc_lattice_colouring": [_P, _I32, _P],
    "csmc_lattice_update_couplings": [_P, _I32, _P],
    "csmc_batch_create": [_P, _I32, _P, _I32, _P],
    "csmc_batch_destroy": [_P, _I32],
    "csmc_set_spins": [_P, _I32, _I32, _P],
    "csmc_get_spins": [_P, _I32, _I32, _P],
    "csmc_randomize_spins": [_P, _I32, _U64],
    "csmc_local_field": [_P, _I32, _I32, _P],
    "csmc_site_energy": [_P, _I32, _I32, _P],
    "csmc_total_energy": [_P, _I32, _P, _P],
    "csmc_magnetization": [_P, _I32, _P],
    "csmc_sweep": [_P, _I32, _I32, _P, _P, _I32, _U64, _P],
    "csmc_sweep_plain": [_P, _I32, _I32, _P, _P, _I32, _U64, _P],
    "csmc_set_counters": [_P, _I32, _P],
    "csmc_get_counters": [_P, _I32, _P],
    "csmc_batch_storage_mode": [_P, _I32, _P],
    "csmc_anneal": [_P, _I32, _P, _I32, _I64, _I64, _I32, _D, _U64, _P],
    "csmc_pt_init": [_P, _I32, _P, _P, _D, _U64],
    "csmc_pt_advance": [_P, _I32, _I64, _P, _P],
    "csmc_pt_state": [_P, _I32, _P, _P, _P, _P, _P],
    "csmc_pt_split_parts": [_P, _I32, _P],
    "csmc_pt_observables": [_P, _I32, _P],
    "csmc_spin_correlations": [_P, _I32, _P, _I32, _P],
    "csmc_pt_set_correlations": [_P, _I32, _P, _I32],
    "csmc_sweep_passes": [_P, _I32, _I32, _P, _P, _I32, C.c_uint64, _P],
    "csmc_autotune": [_P, _I32, _I32, _P, _I32, _P, _P],
    "csmc_graph_sweeps": [_P, _I32, _I32, _P, _P, _I32, C.c_uint64, _P],
    "csmc_graph_launch": [_P, _P],
    "csmc_graph_accepted": [_P, _P, _P, _P],
    "csmc_graph_destroy": [_P],
    "csmc_pt_save": [_P, _I32, _P, _P],
    "csmc_pt_save_async": [_P, _I32],
    "csmc_pt_save_wait": [_P, _I32, _P, _P],
    "csmc_pt_load": [_P, _I32, _P, C.c_int64],
    "csmc_pt_correlations": [_P, _I32, _P, _P],
    "csmc_energy_grad_spins": [_P, _I32, _P],
    "csmc_energy_grad_couplings": [_P, _I32, _P],
    "csmc_energy_grad_couplings_noalloc": [_P, _I32, _P],
    "csmc_gradient_descent": [_P, _I32, _D, _I32],
    "csmc_constraint_set": [_P, _I32, _I32, _P, _P, _D, _D, _D],
    "csmc_constraint_lambda": [_P, _I32, _P, _P],
}

_lib = None


def lib():
    """Load libcsmc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C classicalspinmc.jl_amd` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int32
    L.csmc_last_error.argtypes = []
    L.csmc_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc):
    if rc == OK:
        return
    msg = lib().csmc_last_error().decode(errors="replace")
    if rc == EBOUNDS:
        raise BoundsError(rc, msg)
    if rc == EUNDEF:
        raise UndefError(rc, msg)
    if rc == EBC:
        raise CsmcError(rc, msg)
    raise CsmcError(rc, msg)


def ptr(a):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def device_count():
    n = C.c_int32(0)
    check(lib().csmc_device_count(C.byref(n)))
    return n.value


# Live contexts and captured graphs.  They are released by an atexit handler,
# which runs at the start of interpreter shutdown: their device buffers,
# page-locked host buffers, rocBLAS handle and stream then go back to the HIP
# runtime while it is fully up, never from a finaliser that runs in whatever
# order garbage collection reaches it, after other libraries' exit handlers.
_live_contexts = weakref.WeakSet()
_live_graphs = weakref.WeakSet()


def _shutdown():
    global _default_ctx
    for g in list(_live_graphs):
        g.close()
    for c in list(_live_contexts):
        c.close()
    _default_ctx = None


atexit.register(_shutdown)


def register_graph(g):
    _live_graphs.add(g)


class Context:
    """One libcsmc context (HIP stream + table pools) on one device."""

    _serial = itertools.count(1)

    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().csmc_ctx_create(device, C.byref(self.h)))
        self.device = device
        # cache key for per-context device objects: a handle value can be
        # reused by a later context once this one is destroyed
        self.uid = next(Context._serial)
        _live_contexts.add(self)

    def close(self):
        """Wait for the context's stream and release everything it holds
        (batches, lattice pools, stream).  Batches of the context are
        unusable afterwards.  Idempotent."""
        if self.h:
            for g in list(_live_graphs):
                if g.batch.ctx is self:
                    g.close()
            lib().csmc_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def sync(self):
        check(lib().csmc_ctx_sync(self.h))

    def timing(self, enable=True):
        """start (or stop) a window of HIP-event kernel timing"""
        check(lib().csmc_timing(self.h, 1 if enable else 0))

    def timing_read(self):
        """(summed kernel ms, launches) of the current timing window"""
        ms = C.c_double()
        n = C.c_int64()
        check(lib().csmc_timing_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    """The process-wide context on device 0 (created on first use)."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def as_f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def as_i32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    if shape is not None:
        a = a.reshape(shape)
    return a
