"""SpinBatch: a batch of systems (replicas of one lattice, or independent
lattices of uneven size) resident on one GPU, driven through libcsmc.

This is what replaces one MPI rank per replica (monte_carlo.jl:246-262):
every system of the batch sweeps in the same persistent launch, and
parallel tempering exchanges temperature labels on the device.
"""
import ctypes as C

import numpy as np

from . import _lib


class SpinBatch:
    def __init__(self, lattices, precision=32, ctx=None):
        self.ctx = ctx or _lib.default_context()
        self.lattices = list(lattices)
        if not self.lattices:
            raise ValueError("empty batch")
        self.precision = int(precision)
        ids = np.array([lat.device_id(self.ctx) for lat in self.lattices], dtype=np.int32)
        bid = C.c_int32()
        _lib.check(_lib.lib().csmc_batch_create(self.ctx.h, len(ids), _lib.ptr(ids), self.precision,
                                                C.byref(bid)))
        self.id = bid.value
        self.n = len(ids)
        self.sizes = np.array([lat.size for lat in self.lattices], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])

    @property
    def _h(self):
        return self.ctx.h

    def close(self):
        if self.id is not None and self.ctx.h:
            _lib.lib().csmc_batch_destroy(self.ctx.h, self.id)
        self.id = None

    # -- spins ---------------------------------------------------------------------
    def set_spins(self, i, spins):
        s = _lib.as_f64(spins)
        if s.shape == (3, self.sizes[i]):
            s = np.ascontiguousarray(s.T)
        _lib.check(_lib.lib().csmc_set_spins(self._h, self.id, i, _lib.ptr(s.reshape(-1, 3))))

    def get_spins(self, i):
        out = np.zeros((int(self.sizes[i]), 3))
        _lib.check(_lib.lib().csmc_get_spins(self._h, self.id, i, _lib.ptr(out)))
        return out

    def load_lattices(self):
        for i, lat in enumerate(self.lattices):
            self.set_spins(i, lat._spins)

    def randomize(self, seed):
        _lib.check(_lib.lib().csmc_randomize_spins(self._h, self.id, int(seed)))

    # -- evaluation ---------------------------------------------------------------
    def total_energy(self, parts=False):
        E = np.zeros(self.n)
        P = np.zeros((self.n, 5))
        _lib.check(_lib.lib().csmc_total_energy(self._h, self.id, _lib.ptr(E), _lib.ptr(P)))
        return (E, P) if parts else E

    def magnetization(self):
        M = np.zeros((self.n, 3))
        _lib.check(_lib.lib().csmc_magnetization(self._h, self.id, _lib.ptr(M)))
        return M

    def local_field(self, i):
        H = np.zeros((int(self.sizes[i]), 3))
        _lib.check(_lib.lib().csmc_local_field(self._h, self.id, i, _lib.ptr(H)))
        return H

    def site_energy(self, i):
        e = np.zeros(int(self.sizes[i]))
        _lib.check(_lib.lib().csmc_site_energy(self._h, self.id, i, _lib.ptr(e)))
        return e

    # -- sweeps ---------------------------------------------------------------------
    def _vec(self, v, default=None):
        if v is None:
            return None if default is None else np.full(self.n, float(default))
        return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (self.n,)).copy())

    def sweep(self, kind, T=None, sigma=None, n_sweeps=1, seed=0, plain=False):
        """n_sweeps coloured sweeps of `kind`; returns accepted counts (and
        updates `sigma` in place for the adaptive kind if an array is given)."""
        Tv = self._vec(T)
        sg = sigma if isinstance(sigma, np.ndarray) and sigma.dtype == np.float64 and sigma.shape == (self.n,) \
            else self._vec(sigma)
        acc = np.zeros(self.n)
        fn = _lib.lib().csmc_sweep_plain if plain else _lib.lib().csmc_sweep
        _lib.check(fn(self._h, self.id, int(kind), _lib.ptr(Tv), _lib.ptr(sg), int(n_sweeps), int(seed),
                      _lib.ptr(acc)))
        return acc

    def storage_mode(self):
        """storage mode of the persistent sweep kernel for this batch: 3
        register tables + LDS spins, 2 LDS spins and neighbour tables, 1 LDS
        spins, 0 global memory"""
        m = C.c_int32()
        _lib.check(_lib.lib().csmc_batch_storage_mode(self._h, self.id, C.byref(m)))
        return m.value

    def counters(self):
        c = np.zeros(self.n, dtype=np.uint32)
        _lib.check(_lib.lib().csmc_get_counters(self._h, self.id, _lib.ptr(c)))
        return c

    def set_counters(self, c):
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.uint32), (self.n,)).copy())
        _lib.check(_lib.lib().csmc_set_counters(self._h, self.id, _lib.ptr(c)))

    def set_constraint(self, a, Q, c0, weight, sigma0=60.0):
        """the site-sum constraint c = sum_i (a_b . s_i + s_i . Q_b s_i) - c0
        of metropolis_constraint_adaptive! (csmc_constraint_set); a:
        (n_basis, 3), Q: (n_basis, 3, 3)"""
        a = _lib.as_f64(a).reshape(-1, 3)
        Q = _lib.as_f64(Q).reshape(-1, 9)
        if len(a) != len(Q):
            raise ValueError("one a and one Q per basis")
        _lib.check(_lib.lib().csmc_constraint_set(self._h, self.id, len(a), _lib.ptr(a), _lib.ptr(Q),
                                                  float(c0), float(weight), float(sigma0)))

    def constraint_lambda(self, set_to=None):
        """MonteCarlo.lambda per system (set first when set_to is given)"""
        s = None if set_to is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(set_to, dtype=np.float64), (self.n,)).copy())
        out = np.zeros(self.n)
        _lib.check(_lib.lib().csmc_constraint_lambda(self._h, self.id, _lib.ptr(s), _lib.ptr(out)))
        return out

    def anneal(self, temps, t_thermalization, overrelaxation_rate, alg, sigma0=60.0, seed=0):
        temps = _lib.as_f64(temps).reshape(-1)
        acc = np.zeros((self.n, len(temps)))
        _lib.check(_lib.lib().csmc_anneal(self._h, self.id, _lib.ptr(temps), len(temps),
                                          int(t_thermalization), int(overrelaxation_rate), int(alg),
                                          float(sigma0), int(seed), _lib.ptr(acc)))
        return acc

    # -- parallel tempering -------------------------------------------------------
    def pt_init(self, temps, params, alg, sigma0=60.0, seed=0, measure=True):
        temps = _lib.as_f64(temps).reshape(-1)
        if temps.shape != (self.n,):
            raise ValueError("one temperature per system")
        p = _lib.PTParams()
        for k in ("t_thermalization", "t_measurement", "probe_rate", "swap_rate",
                  "overrelaxation_rate", "report_interval", "checkpoint_rate"):
            setattr(p, k, int(params[k]))
        p.alg = int(alg)
        p.measure = 1 if measure else 0
        self.temps = temps
        _lib.check(_lib.lib().csmc_pt_init(self._h, self.id, _lib.ptr(temps), C.byref(p), float(sigma0),
                                           int(seed)))

    def pt_advance(self, until=-1):
        stopped = C.c_int64()
        event = C.c_int32()
        _lib.check(_lib.lib().csmc_pt_advance(self._h, self.id, int(until), C.byref(stopped),
                                              C.byref(event)))
        return stopped.value, event.value

    def pt_state(self):
        sos = np.zeros(self.n, dtype=np.int32)
        E = np.zeros(self.n)
        al = np.zeros(self.n)
        ax = np.zeros(self.n)
        sg = np.zeros(self.n)
        _lib.check(_lib.lib().csmc_pt_state(self._h, self.id, _lib.ptr(sos), _lib.ptr(E), _lib.ptr(al),
                                            _lib.ptr(ax), _lib.ptr(sg)))
        return dict(system_of_slot=sos, E=E, accepted_local=al, accepted_exchange=ax, sigma=sg)

    def spin_correlations(self, ks):
        """compute_equal_time_correlations (spin_correlations.jl:6-43) of every
        system (they must share one lattice): (n, 9, n_k), FP64."""
        ks = np.asarray(ks, dtype=np.float64)
        if ks.ndim != 2:
            raise ValueError("ks must be a D x n_k matrix")
        kt = np.ascontiguousarray(ks.T)
        out = np.zeros((self.n, ks.shape[1], 9))
        _lib.check(_lib.lib().csmc_spin_correlations(self._h, self.id, _lib.ptr(kt), ks.shape[1],
                                                     _lib.ptr(out)))
        return np.transpose(out, (0, 2, 1)).copy()

    def pt_set_correlations(self, ks):
        """measure compute_equal_time_correlations at the momenta ks (D x n_k)
        at every probe sweep of the PT run (corr = true,
        monte_carlo.jl:371-375); call after pt_init"""
        ks = np.asarray(ks, dtype=np.float64)
        if ks.ndim != 2:
            raise ValueError("ks must be a D x n_k matrix")
        kt = np.ascontiguousarray(ks.T)
        _lib.check(_lib.lib().csmc_pt_set_correlations(self._h, self.id, _lib.ptr(kt), ks.shape[1]))
        self._corr_nk = ks.shape[1]

    def pt_correlations(self):
        """(mean over probes per temperature slot, (n, 9, n_k); probe count)"""
        nk = getattr(self, "_corr_nk", 0)
        out = np.zeros((self.n, max(nk, 1), 9))
        cnt = C.c_int64()
        _lib.check(_lib.lib().csmc_pt_correlations(self._h, self.id, _lib.ptr(out), C.byref(cnt)))
        return np.transpose(out, (0, 2, 1)).copy(), cnt.value

    def sweep_passes(self, kind, T=None, sigma=None, n_sweeps=1, seed=0):
        """the plain form over the whole batch (csmc_sweep_passes): one
        launch per colour pass covering every system; returns accepted
        counts"""
        Tv = self._vec(T)
        sg = self._vec(sigma)
        acc = np.zeros(self.n)
        _lib.check(_lib.lib().csmc_sweep_passes(self._h, self.id, int(kind), _lib.ptr(Tv), _lib.ptr(sg),
                                                int(n_sweeps), C.c_uint64(int(seed)), _lib.ptr(acc)))
        return acc

    def tiles(self):
        """tiles per system of the batch-wide sweeps (0: one launch per
        colour pass) and, for 0, the reason"""
        n = C.c_int32(0)
        _lib.check(_lib.lib().csmc_batch_tiles(self._h, self.id, C.byref(n), None))
        return n.value, ("" if n.value else _lib.lib().csmc_last_error().decode())

    def tile_sweeps(self):
        """sweeps one tiled launch runs while that many remain (1: one
        launch per sweep, or no tiles)"""
        n, k = C.c_int32(0), C.c_int32(1)
        _lib.check(_lib.lib().csmc_batch_tiles(self._h, self.id, C.byref(n), C.byref(k)))
        return k.value

    AUTOTUNE_FORMS = ("persistent, storage mode cap 3", "persistent, cap 2", "persistent, cap 1",
                      "persistent, cap 0", "batch-wide colour passes")

    def autotune(self, kind=None, T=None, n_sweeps=3):
        """time each sweep form on the batch's own state and keep the
        fastest for later sweeps (csmc_autotune); returns {form: ms} and the
        chosen form"""
        kind = _lib.OVERRELAX if kind is None else kind
        Tv = self._vec(T)
        times = np.zeros(5)
        choice = C.c_int32()
        _lib.check(_lib.lib().csmc_autotune(self._h, self.id, int(kind), _lib.ptr(Tv), int(n_sweeps),
                                            _lib.ptr(times), C.byref(choice)))
        return ({f: t for f, t in zip(self.AUTOTUNE_FORMS, times) if np.isfinite(t)},
                self.AUTOTUNE_FORMS[choice.value])

    def graph_sweeps(self, kind, T=None, sigma=None, n_sweeps=1, seed=0):
        """n_sweeps plain-form sweeps captured as one hipGraph
        (csmc_graph_sweeps); replay with .launch()"""
        Tv = self._vec(T) if T is not None else None
        sg = self._vec(sigma) if sigma is not None else None
        g = C.c_void_p()
        _lib.check(_lib.lib().csmc_graph_sweeps(self._h, self.id, int(kind), _lib.ptr(Tv), _lib.ptr(sg),
                                                int(n_sweeps), C.c_uint64(int(seed)), C.byref(g)))
        return SweepGraph(self, g)

    def pt_save(self):
        """the whole PT state as bytes (csmc_pt_save)"""
        n = C.c_int64()
        _lib.check(_lib.lib().csmc_pt_save(self._h, self.id, None, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        _lib.check(_lib.lib().csmc_pt_save(self._h, self.id, _lib.ptr(buf), C.byref(n)))
        return buf[:n.value].tobytes()

    def pt_save_async(self):
        """start a save that overlaps the following pt_advance calls"""
        _lib.check(_lib.lib().csmc_pt_save_async(self._h, self.id))

    def pt_save_wait(self):
        """the state captured by pt_save_async, as bytes"""
        n = C.c_int64()
        _lib.check(_lib.lib().csmc_pt_save_wait(self._h, self.id, None, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        _lib.check(_lib.lib().csmc_pt_save_wait(self._h, self.id, _lib.ptr(buf), C.byref(n)))
        return buf[:n.value].tobytes()

    def pt_load(self, state):
        """continue a saved run (after pt_init with the saved run's settings)"""
        buf = np.frombuffer(bytes(state), dtype=np.uint8).copy()
        _lib.check(_lib.lib().csmc_pt_load(self._h, self.id, _lib.ptr(buf), len(buf)))

    def pt_split_parts(self):
        """workgroups per system of the chained PT kernel (1: whole system
        per workgroup; G > 1: strips with halos)"""
        g = C.c_int32()
        _lib.check(_lib.lib().csmc_pt_split_parts(self._h, self.id, C.byref(g)))
        return g.value

    def pt_observables(self):
        out = np.zeros((self.n, 8))
        _lib.check(_lib.lib().csmc_pt_observables(self._h, self.id, _lib.ptr(out)))
        keys = ("specific_heat", "specific_heat_err", "susceptibility", "susceptibility_err",
                "energy", "energy_err", "magnetization", "magnetization_err")
        return {k: out[:, i].copy() for i, k in enumerate(keys)}

    # -- adjoint --------------------------------------------------------------------
    def grad_spins(self):
        g = np.zeros((int(self.offsets[-1]), 3))
        _lib.check(_lib.lib().csmc_energy_grad_spins(self._h, self.id, _lib.ptr(g)))
        return [g[self.offsets[i]:self.offsets[i + 1]] for i in range(self.n)]

    def grad_couplings(self):
        n = self.lattices[0].unit_cell.n_params()
        g = np.zeros((self.n, n))
        _lib.check(_lib.lib().csmc_energy_grad_couplings(self._h, self.id, _lib.ptr(g)))
        return g

    def gradient_descent(self, eta, n_steps=1):
        _lib.check(_lib.lib().csmc_gradient_descent(self._h, self.id, float(eta), int(n_steps)))

    def sync(self):
        _lib.check(_lib.lib().csmc_ctx_sync(self._h))


class SweepGraph:
    """A captured block of plain-form sweeps (csmc_graph_*)."""

    def __init__(self, batch, handle):
        self.batch = batch
        self.h = handle
        _lib.register_graph(self)

    def launch(self):
        """enqueue one replay (returns without waiting)"""
        _lib.check(_lib.lib().csmc_graph_launch(self.batch._h, self.h))

    def accepted(self):
        """(accepted proposals per system of the last replay, cone widths)"""
        acc = np.zeros(self.batch.n)
        sg = np.zeros(self.batch.n)
        _lib.check(_lib.lib().csmc_graph_accepted(self.batch._h, self.h, _lib.ptr(acc), _lib.ptr(sg)))
        return acc, sg

    def close(self):
        if self.h:
            _lib.lib().csmc_graph_destroy(self.h)
            self.h = None
