"""The HDF5 file layout of ClassicalSpinMC.jl (hdf5.jl:1-270).

Files:
  <outpath><prefix>.h5.params   written by rank 0 (monte_carlo.jl:99-107):
      root attributes: the SimulationParameters fields + user inparams
      unit_cell/{lattice_vectors, basis, field/<b>, onsite/<b>,
                 bilinear/"(b1,b2),(o..)", cubic/..., quartic/...}
      lattice/{size, S, bc}
  <outpath><prefix>_<rank>.h5   one per temperature slot (monte_carlo.jl:96):
      attributes T, paramsfile; datasets spins, site_positions;
      group observables/{specific_heat(_err), susceptibility(_err),
                         magnetization(_err), energy(_err)}
  <dir>/IC_<rank>/IC_<k>.h5     attributes T, paramsfile; dataset spins

Julia writes arrays column-major, so every dataset appears on disk with its
Julia dimensions reversed: spins (3 x N in Julia) is an (N, 3) dataset.  The
arrays below are written in that on-disk (C) order.
"""
import ast
import ctypes as C
import os

import numpy as np

from .lattice import Lattice
from .unitcell import (UnitCell, addBasisSite, addBilinear, addCubic, addOnSite, addQuartic,
                       addZeemanCoupling)

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libcsmc_h5.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `make -C classicalspinmc.jl_amd`")
        L = C.CDLL(path)
        I64, P, S = C.c_int64, C.c_void_p, C.c_char_p
        sig = {
            "h5c_open": ([S, C.c_int], I64), "h5c_close": ([I64], C.c_int),
            "h5c_exists": ([I64, S], C.c_int), "h5c_group": ([I64, S, C.c_int], I64),
            "h5c_gclose": ([I64], C.c_int), "h5c_delete": ([I64, S], C.c_int),
            "h5c_write_f64": ([I64, S, C.c_int, P, P], C.c_int),
            "h5c_write_i64": ([I64, S, C.c_int, P, P], C.c_int),
            "h5c_write_u8": ([I64, S, C.c_int, P, P], C.c_int),
            "h5c_read_u8": ([I64, S, P], C.c_int),
            "h5c_write_str": ([I64, S, S], C.c_int),
            "h5c_attr_f64": ([I64, S, C.c_double], C.c_int),
            "h5c_attr_i64": ([I64, S, I64], C.c_int),
            "h5c_attr_str": ([I64, S, S], C.c_int),
            "h5c_info": ([I64, S, P, P, P], C.c_int),
            "h5c_read_f64": ([I64, S, P], C.c_int), "h5c_read_i64": ([I64, S, P], C.c_int),
            "h5c_read_str": ([I64, S, P, C.c_int], C.c_int),
            "h5c_attr_read": ([I64, S, P, P, P, P, C.c_int], C.c_int),
            "h5c_list": ([I64, C.c_int, P, C.c_int], C.c_int),
        }
        for n, (a, r) in sig.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def _b(s):
    return s.encode()


def _chk(rc, what):
    if rc < 0:
        raise OSError(f"HDF5 operation failed: {what}")
    return rc


class H5Node:
    """a file or group handle"""

    def __init__(self, hid, is_file=False):
        self.hid = hid
        self.is_file = is_file

    def close(self):
        if self.hid >= 0:
            (lib().h5c_close if self.is_file else lib().h5c_gclose)(self.hid)
            self.hid = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- children ---------------------------------------------------------------
    def group(self, name, create=False):
        if create and not self.has(name):
            return H5Node(_chk(lib().h5c_group(self.hid, _b(name), 1), f"create group {name}"))
        return H5Node(_chk(lib().h5c_group(self.hid, _b(name), 0), f"open group {name}"))

    def has(self, name):
        return lib().h5c_exists(self.hid, _b(name)) == 1

    def delete(self, name):
        _chk(lib().h5c_delete(self.hid, _b(name)), f"delete {name}")

    def keys(self):
        buf = C.create_string_buffer(1 << 20)
        _chk(lib().h5c_list(self.hid, 0, buf, len(buf)), "list")
        return [k for k in buf.value.decode().split("\n") if k]

    def attr_keys(self):
        buf = C.create_string_buffer(1 << 16)
        _chk(lib().h5c_list(self.hid, 1, buf, len(buf)), "list attributes")
        return [k for k in buf.value.decode().split("\n") if k]

    # -- datasets ---------------------------------------------------------------
    def write(self, name, value):
        if isinstance(value, str):
            _chk(lib().h5c_write_str(self.hid, _b(name), _b(value)), name)
            return
        a = np.asarray(value)
        shape = a.shape
        dims = np.array(list(shape) + [0], dtype=np.int64)  # never empty
        if a.dtype == np.uint8:
            a = np.ascontiguousarray(a)
            fn = lib().h5c_write_u8
        elif np.issubdtype(a.dtype, np.integer) or a.dtype == bool:
            a = np.ascontiguousarray(a, dtype=np.int64)  # (ascontiguousarray makes 0-d 1-d)
            fn = lib().h5c_write_i64
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
            fn = lib().h5c_write_f64
        _chk(fn(self.hid, _b(name), len(shape), dims.ctypes.data, a.ctypes.data), name)

    def read(self, name):
        kind, rank = C.c_int(), C.c_int()
        dims = np.zeros(8, dtype=np.int64)
        _chk(lib().h5c_info(self.hid, _b(name), C.byref(kind), C.byref(rank), dims.ctypes.data),
             f"no dataset {name}")
        if kind.value == 2:
            n = int(dims[0]) if dims[0] > 0 else 1 << 16
            buf = C.create_string_buffer(n + 1)
            _chk(lib().h5c_read_str(self.hid, _b(name), buf, n + 1), name)
            return buf.value.decode()
        shape = tuple(int(d) for d in dims[: rank.value])
        if kind.value == 3:
            out = np.zeros(shape, dtype=np.uint8)
            _chk(lib().h5c_read_u8(self.hid, _b(name), out.ctypes.data), name)
        elif kind.value == 1:
            out = np.zeros(shape, dtype=np.int64)
            _chk(lib().h5c_read_i64(self.hid, _b(name), out.ctypes.data), name)
        else:
            out = np.zeros(shape, dtype=np.float64)
            _chk(lib().h5c_read_f64(self.hid, _b(name), out.ctypes.data), name)
        return out if rank.value else out.reshape(())[()]

    # -- attributes -------------------------------------------------------------
    def set_attr(self, name, value):
        if isinstance(value, str):
            _chk(lib().h5c_attr_str(self.hid, _b(name), _b(value)), name)
        elif isinstance(value, (bool, int, np.integer)):
            _chk(lib().h5c_attr_i64(self.hid, _b(name), int(value)), name)
        else:
            _chk(lib().h5c_attr_f64(self.hid, _b(name), float(value)), name)

    def get_attr(self, name):
        kind = C.c_int()
        f = C.c_double()
        i = C.c_int64()
        buf = C.create_string_buffer(1 << 16)
        _chk(lib().h5c_attr_read(self.hid, _b(name), C.byref(kind), C.byref(f), C.byref(i), buf,
                                 len(buf)), f"no attribute {name}")
        return (f.value, i.value, buf.value.decode())[kind.value]


def h5open(path, mode="r"):
    m = {"r": 0, "r+": 1, "w": 2}[mode]
    return H5Node(_chk(lib().h5c_open(_b(path), m), f"open {path} ({mode})"), is_file=True)


def _jl_tuple(t):
    """Julia's printing of an Int tuple: (0, -1), (1,), (0, 0, 0)"""
    t = tuple(int(v) for v in t)
    return "(" + str(t[0]) + ",)" if len(t) == 1 else "(" + ", ".join(str(v) for v in t) + ")"


# -- hdf5.jl functions ---------------------------------------------------------

def dump_attributes_hdf5(node, obj):
    """dump_attributes_hdf5! (hdf5.jl:6-22)"""
    items = obj.items() if isinstance(obj, dict) else obj.as_dict().items()
    for k, v in items:
        node.set_attr(k, v)


def write_attributes(filename, d):
    """write_attributes (hdf5.jl:27-31)"""
    with h5open(filename, "r+") as f:
        dump_attributes_hdf5(f, d)


def dump_unit_cell(fid, uc):
    """dump_unit_cell! (hdf5.jl:36-76)"""
    with fid.group("unit_cell", create=True) as g:
        g.write("lattice_vectors", np.stack(uc.lattice_vectors))  # rows are a_i on disk
        g.write("basis", np.array(uc.basis).T)                       # (D, nb) on disk
        with g.group("field", create=True) as h:
            for b, vec in uc.field:
                h.write(str(b), vec)
        with g.group("onsite", create=True) as o:
            for b, M in uc.onsite:
                o.write(str(b), np.asarray(M).T)
        with g.group("bilinear", create=True) as bl:
            for b1, b2, M, off in uc.bilinear:
                bl.write(f"({b1},{b2}),{_jl_tuple(off)}", np.asarray(M).T)
        with g.group("cubic", create=True) as c:
            for b1, b2, b3, M, o2, o3 in uc.cubic:
                c.write(f"({b1},{b2},{b3}),{_jl_tuple(o2)},{_jl_tuple(o3)}", np.asarray(M).transpose(2, 1, 0))
        with g.group("quartic", create=True) as q:
            for b1, b2, b3, b4, M, o2, o3, o4 in uc.quartic:
                q.write(f"({b1},{b2},{b3},{b4}),{_jl_tuple(o2)},{_jl_tuple(o3)},{_jl_tuple(o4)}",
                        np.asarray(M).transpose(3, 2, 1, 0))


def read_unit_cell(fid):
    """read_unit_cell (hdf5.jl:81-120); keys are parsed with ast.literal_eval"""
    with fid.group("unit_cell") as g:
        lv = g.read("lattice_vectors")
        uc = UnitCell(*[lv[i] for i in range(lv.shape[0])])
        basis = g.read("basis").T
        for site in basis:
            addBasisSite(uc, site)
        with g.group("field") as h:
            for k in h.keys():
                addZeemanCoupling(uc, int(k), h.read(k))
        with g.group("onsite") as o:
            for k in o.keys():
                addOnSite(uc, int(k), o.read(k).T)
        with g.group("bilinear") as bl:
            for k in bl.keys():
                bs, off = ast.literal_eval(k)
                addBilinear(uc, bs[0], bs[1], bl.read(k).T, off)
        with g.group("cubic") as c:
            for k in c.keys():
                bs, o2, o3 = ast.literal_eval(k)
                addCubic(uc, *bs, c.read(k).transpose(2, 1, 0), o2, o3)
        with g.group("quartic") as q:
            for k in q.keys():
                bs, o2, o3, o4 = ast.literal_eval(k)
                addQuartic(uc, *bs, q.read(k).transpose(3, 2, 1, 0), o2, o3, o4)
    return uc


def dump_metadata(fid, mc):
    """dump_metadata! (hdf5.jl:125-137)"""
    dump_unit_cell(fid, mc.lattice.unit_cell)
    with fid.group("lattice", create=True) as l:
        l.write("size", np.array(mc.lattice.shape, dtype=np.int64))
        l.write("S", float(mc.lattice.S))
        l.write("bc", mc.lattice.bc)
    dump_attributes_hdf5(fid, mc.parameters)


def create_params_file(mc, filename):
    """create_params_file (hdf5.jl:142-147)"""
    with h5open(filename, "w") as f:
        dump_metadata(f, mc)
    return filename


def read_lattice(fid, **kw):
    """read_lattice (hdf5.jl:152-159)"""
    with fid.group("lattice") as l:
        size = tuple(int(v) for v in l.read("size"))
        S = float(l.read("S"))
        bc = l.read("bc")
    uc = read_unit_cell(fid)
    return Lattice(size, uc, S, bc=bc, **kw)


def initialize_hdf5(mc, paramsfile):
    """initialize_hdf5 (hdf5.jl:164-171)"""
    with h5open(mc.outpath, "w") as f:
        f.set_attr("T", float(mc.T))
        f.set_attr("paramsfile", paramsfile)
        f.write("spins", mc.lattice._spins)                    # (N, 3) on disk
        f.write("site_positions", mc.lattice.site_positions.T)  # (N, D) on disk


def write_MC_checkpoint(mc):
    """write_MC_checkpoint (hdf5.jl:176-180): overwrite the spins dataset"""
    with h5open(mc.outpath, "r+") as f:
        f.write("spins", mc.lattice._spins)


def write_resume_state(mc, state, sweep):
    """the saved parallel-tempering state of the whole run (csmc_pt_save) in
    the group `resume` of mc's file: `state` (bytes) and `sweep`; no
    counterpart in the reference, whose checkpoints hold the spins only"""
    with h5open(mc.outpath, "r+") as f:
        with f.group("resume", create=True) as g:
            g.write("state", np.frombuffer(state, dtype=np.uint8))
            g.write("sweep", int(sweep))


def read_resume_state(mc):
    """(state bytes, sweep) from mc's file, or None"""
    if not mc.outpath or not os.path.isfile(mc.outpath):
        return None
    with h5open(mc.outpath, "r") as f:
        if not f.has("resume"):
            return None
        with f.group("resume") as g:
            return g.read("state").tobytes(), int(g.read("sweep"))


def write_initial_configuration(filename, mc):
    """write_initial_configuration (hdf5.jl:185-194)"""
    with h5open(mc.outpath, "r") as p:
        paramsfile = p.get_attr("paramsfile")
    with h5open(filename, "w") as f:
        f.set_attr("T", float(mc.T))
        f.set_attr("paramsfile", paramsfile)
        f.write("spins", mc.lattice._spins)


def ic_path(mc, rank, k):
    d = os.path.join(os.path.dirname(mc.outpath), f"IC_{rank}")
    os.makedirs(d, exist_ok=True)
    return os.path.join(d, f"IC_{k}.h5")


def write_final_observables(mc):
    """write_final_observables (hdf5.jl:204-239): C and chi errors are written
    relative (|dC/C|), E and |M| errors absolute, as the reference does."""
    o = mc.observables
    with h5open(mc.outpath, "r+") as f:
        f.write("spins", mc.lattice._spins)
        if f.has("observables"):
            with f.group("observables") as g:
                for k in g.keys():
                    g.delete(k)
            print(f"Overwriting observables in {mc.outpath}")
        with f.group("observables", create=True) as g:
            g.write("specific_heat", o["specific_heat"])
            g.write("specific_heat_err", abs(o["specific_heat_err"] / o["specific_heat"])
                    if o["specific_heat"] else float("nan"))
            g.write("susceptibility", o["susceptibility"])
            g.write("susceptibility_err", abs(o["susceptibility_err"] / o["susceptibility"])
                    if o["susceptibility"] else float("nan"))
            g.write("magnetization", o["magnetization"])
            g.write("magnetization_err", o["magnetization_err"])
            g.write("energy", o["energy"])
            g.write("energy_err", o["energy_err"])
        if mc.corr:
            # Julia's 9 x n_k and D x n_k matrices, column-major, are the
            # row-major n_k x 9 and n_k x D arrays written here
            with f.group("spin_correlations", create=True) as g:
                overwrite_keys(g, {"SSF": np.ascontiguousarray(np.asarray(mc.correlations).T),
                                   "SSF_momentum": np.ascontiguousarray(np.asarray(mc.momentum_vectors).T)})
            print(f"Writing spin correlations to {mc.outpath}")


def overwrite_keys(fid, d):
    """overwrite_keys! (hdf5.jl:244-252)"""
    for k, v in d.items():
        if fid.has(k):
            fid.delete(k)
            print(f"Overwriting parameter {k}")
        fid.write(k, v)


def attributes_to_dict(f):
    """attributes_to_dict (hdf5.jl:254-261)"""
    return {k: f.get_attr(k) for k in f.attr_keys()}


def read_spin_configuration(lat, filename):
    """read_spin_configuration! (hdf5.jl:266-270)"""
    with h5open(filename, "r") as f:
        lat._spins[:] = f.read("spins")


def setup_output(mc, outpath, outprefix, inparams, overwrite, rank):
    """file creation part of the MonteCarlo constructor (monte_carlo.jl:94-117)"""
    filename = f"{outprefix}_{rank}.h5"
    if rank == 0 and not os.path.isdir(outpath):
        os.makedirs(outpath, exist_ok=True)
    mc.outpath = outpath + filename
    paramsfile = outpath + outprefix + ".h5.params"
    if rank == 0 and not os.path.isfile(paramsfile) and overwrite:
        create_params_file(mc, paramsfile)
        if inparams:
            write_attributes(paramsfile, inparams)
    if not os.path.isfile(mc.outpath) and overwrite:
        print(f"Creating new file {filename} for output on rank {rank}")
        initialize_hdf5(mc, paramsfile)
