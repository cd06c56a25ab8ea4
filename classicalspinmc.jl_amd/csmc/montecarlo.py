"""MonteCarlo, the Metropolis `alg` plugins and the drivers
simulated_annealing / deterministic_updates / parallel_tempering
(metropolis.jl, monte_carlo.jl, helper.jl), running on libcsmc.

Differences from the reference that follow from the GPU design:
  * sweeps visit the colour classes of the interaction hypergraph in turn
    (every site once per sweep) instead of N random sites with replacement;
    within a class this is the reference's own loop (the sites are
    independent), so observables agree statistically, not trajectory-wise;
  * deterministic_updates runs round((t_deterministic - 1) / N) coloured
    sweeps (at least one when t_deterministic > 1) instead of
    t_deterministic - 1 random single-site updates;
  * parallel tempering runs all replicas of a process in one batch and swaps
    temperature labels on the device instead of spins over MPI;
  * random numbers come from Philox4x32-10 keyed by `seed`.
"""
import copy
import datetime
import warnings

import numpy as np

from . import _lib
from .batch import SpinBatch

_KEYS = ["t_thermalization", "t_deterministic", "t_measurement", "probe_rate", "swap_rate",
         "overrelaxation_rate", "report_interval", "checkpoint_rate"]
_DEFAULTS = [1, 1, 1, 1, 1, 10, 0, 0]


class SimulationParameters:
    """SimulationParameters (metropolis.jl:4-13)."""

    def __init__(self, *vals):
        for k, v in zip(_KEYS, vals):
            setattr(self, k, int(v))

    def as_dict(self):
        return {k: getattr(self, k) for k in _KEYS}


def MCParamsBuffer(d):
    """MCParamsBuffer (monte_carlo.jl:10-27): fills defaults INTO the caller's
    dict and warns about unknown keys."""
    vals = []
    for k, dv in zip(_KEYS, _DEFAULTS):
        if k not in d:
            d[k] = dv
        vals.append(d[k])
    for k in d:
        if k not in _KEYS:
            warnings.warn(f"'{k}' not a valid MC parameter; ignoring")
    return SimulationParameters(*vals)


class _Alg:
    """The `alg` plugin (metropolis.jl:181-199): alg(mc, T) performs one
    Metropolis sweep of mc's lattice at T and returns the accepted count."""

    def __init__(self, kind, name):
        self.kind = kind
        self.name = name

    def __call__(self, mc, T):
        b = mc._batch()
        if self.kind == _lib.CONSTRAINT_ADAPTIVE:
            _install_constraint(mc, b)
        sg = np.array([float(mc.sigma)])
        acc = b.sweep(self.kind, T=T, sigma=sg, n_sweeps=1, seed=mc.seed)
        if self.kind in (_lib.CONE_ADAPTIVE, _lib.CONSTRAINT_ADAPTIVE):
            mc.sigma = float(sg[0])
        if self.kind == _lib.CONSTRAINT_ADAPTIVE:
            mc.lambda_ = float(b.constraint_lambda()[0])
        mc._device_dirty = True
        return float(acc[0])

    def __repr__(self):
        return f"{self.name}()"


def Metropolis():
    return _Alg(_lib.METROPOLIS, "Metropolis")


def MetropolisAdaptive():
    return _Alg(_lib.CONE_ADAPTIVE, "MetropolisAdaptive")


def MetropolisFixedCone():
    return _Alg(_lib.CONE, "MetropolisFixedCone")


def _kind(alg):
    """the device sweep kind of an `alg` plugin; a plugin with no device form
    (a user function of (mc, T)) is refused with the reason"""
    if isinstance(alg, _Alg):
        return alg.kind
    raise TypeError(
        f"alg={alg!r} has no device form: the drivers run the alg plugin as a coloured sweep kernel, "
        "so alg must be Metropolis(), MetropolisAdaptive(), MetropolisFixedCone() "
        "(metropolis.jl:181-199) or MetropolisConstraintAdaptive() with a SiteSumConstraint")


class SiteSumConstraint:
    """The device form of MonteCarlo.constraint for
    metropolis_constraint_adaptive! (metropolis.jl:155-179): a sum over sites
    of a per-basis linear and quadratic term,
        c(lattice) = sum_i (a_b . s_i + s_i . Q_b s_i) - c0,  b = basis of i,
    e.g. a fixed magnetisation along n (a_b = n, Q_b = 0, c0 = N m0) or a
    staggered one (a_b = +-n per basis).  Calling it on a Lattice evaluates it
    on the host, as the reference's closure would be called."""

    def __init__(self, a, Q=None, c0=0.0):
        self.a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
        self.Q = (np.zeros((len(self.a), 3, 3)) if Q is None
                  else np.asarray(Q, dtype=np.float64).reshape(-1, 3, 3))
        if len(self.Q) != len(self.a):
            raise ValueError("one a and one Q per basis")
        self.c0 = float(c0)

    def __call__(self, lattice):
        s = np.asarray(lattice._spins)
        b = np.repeat(np.arange(len(lattice.unit_cell.basis)), lattice.size // len(lattice.unit_cell.basis))
        c = 0.0
        for k in range(min(len(self.a), b.max() + 1)):
            sk = s[b == k]
            c += float(np.sum(sk @ self.a[k]) + np.einsum("ia,ab,ib->", sk, self.Q[k], sk))
        return c - self.c0


def _install_constraint(mc, b):
    """hand mc.constraint, mc.weight, mc.sigma0 and mc.lambda to the batch"""
    if not isinstance(mc.constraint, SiteSumConstraint):
        raise TypeError("MetropolisConstraintAdaptive on the device needs mc.constraint to be a "
                        "SiteSumConstraint: the reference calls an arbitrary Julia function per proposal "
                        "(metropolis.jl:164), which a GPU kernel cannot run")
    k = mc.constraint
    b.set_constraint(k.a, k.Q, k.c0, mc.weight, mc.sigma0)
    b.constraint_lambda(mc.lambda_)


def MetropolisConstraint():
    """metropolis_constraint! (metropolis.jl:36-58) calls e_diff, which the
    reference never defines (:44): calling it raises in the reference, so it
    has no device form."""
    raise NotImplementedError("MetropolisConstraint calls the undefined e_diff (metropolis.jl:44); "
                              "use MetropolisConstraintAdaptive")


def MetropolisConstraintAdaptive():
    """metropolis_constraint_adaptive! (metropolis.jl:155-179) for a
    SiteSumConstraint in mc.constraint: runs in csmc_sweep (the plugin call)
    and simulated_annealing; parallel tempering refuses it."""
    return _Alg(_lib.CONSTRAINT_ADAPTIVE, "MetropolisConstraintAdaptive")


class MonteCarlo:
    """MonteCarlo (metropolis.jl:15-29, constructor monte_carlo.jl:56-118).

    Additive keyword arguments: precision (32 or 64, the device state type),
    seed (Philox key), ctx (libcsmc context)."""

    def __init__(self, T, lattice, parameters, constraint=None, weight=0.0, outpath="",
                 outprefix="configuration", inparams=None, overwrite=True, sigma0=60, corr=False,
                 ks=None, precision=64, seed=0, ctx=None, rank=0):
        if corr and (ks is None or np.size(ks) == 0):
            warnings.warn("No momentum vectors provided for correlation calculations!")
        elif not corr and ks is not None and np.size(ks) != 0:
            warnings.warn("Momentum vectors provided but correlation calculations not requested!")
        self.T = float(T)
        self.lattice = copy.deepcopy(lattice)
        self.parameters = MCParamsBuffer(parameters)
        self.observables = None
        self.lambda_ = 0.0
        self.weight = weight
        self.constraint = constraint
        self.sigma = sigma0
        self.sigma0 = sigma0
        self.corr = bool(corr)
        D = lattice.unit_cell.D
        self.momentum_vectors = (np.zeros((D, 0)) if ks is None or np.size(ks) == 0
                                 else np.asarray(ks, dtype=np.float64).reshape(D, -1))
        self.correlations = None
        self.precision = precision
        self.seed = int(seed)
        self.ctx = ctx
        self.rank = rank
        self._b = None
        self._device_dirty = False
        self.outpath = ""
        if outpath:
            from . import hdf5io
            hdf5io.setup_output(self, outpath, outprefix, inparams or {}, overwrite, rank)

    # device batch holding this MonteCarlo's configuration
    def _batch(self):
        if self._b is None:
            self._b = SpinBatch([self.lattice], precision=self.precision, ctx=self.ctx)
            self._b.set_counters(0)
        if not self._device_dirty:
            self._b.set_spins(0, self.lattice._spins)
            self._device_dirty = True
        return self._b

    def sync_host(self):
        """copy the device configuration back into mc.lattice.spins"""
        if self._b is not None and self._device_dirty:
            self.lattice._spins[:] = self._b.get_spins(0)
            self._device_dirty = False


def _schedule_temps(schedule, T0, Tmin):
    temps = []
    T, time = T0, 1
    while T > Tmin:
        temps.append(T)
        T = schedule(time)
        time += 1
    return temps


def simulated_annealing(mc, schedule, T0=1.0, alg=None):
    """simulated_annealing! (monte_carlo.jl:157-190)."""
    alg = alg or Metropolis()
    kind = _kind(alg)
    p = mc.parameters
    N = mc.lattice.size
    accept_total = p.t_thermalization * N
    if p.overrelaxation_rate != 0:
        accept_total /= p.overrelaxation_rate
    b = mc._batch()
    if kind == _lib.CONSTRAINT_ADAPTIVE:
        _install_constraint(mc, b)
    for T in _schedule_temps(schedule, T0, mc.T):
        acc = b.anneal([T], p.t_thermalization, p.overrelaxation_rate, kind, mc.sigma0, mc.seed)
        if kind == _lib.CONSTRAINT_ADAPTIVE:
            mc.lambda_ = float(b.constraint_lambda()[0])
        R = float(acc[0, 0])
        print(f"Acceptance rate at T={T}: {round(R / accept_total * 100, 5)} % ")
        mc.sigma = mc.sigma0
        if mc.outpath:
            from . import hdf5io
            mc.sync_host()
            hdf5io.write_MC_checkpoint(mc)
    mc.sync_host()


def deterministic_updates(mc):
    """deterministic_updates! (monte_carlo.jl:201-213) as coloured sweeps."""
    n_updates = mc.parameters.t_deterministic - 1
    if n_updates <= 0:
        return
    n = max(1, int(round(n_updates / mc.lattice.size)))
    b = mc._batch()
    b.sweep(_lib.DETERMINISTIC, n_sweeps=n, seed=mc.seed)
    mc._device_dirty = True
    mc.sync_host()


def overrelaxation(mc, n_sweeps=1):
    """overrelaxation! (monte_carlo.jl:126-139) on mc's lattice."""
    b = mc._batch()
    b.sweep(_lib.OVERRELAX, n_sweeps=n_sweeps)
    mc._device_dirty = True


def _stats_line(mcs, t, total, stats, prev):
    """print_runtime_statistics! (helper.jl:25-78) from device counters."""
    p = mcs[0].parameters
    progress = 100.0 * t / total
    therm = "YES" if t >= p.t_thermalization else "NO"
    s = f"Sweep {t} / {total} ({progress:.1f}%)\n\t\tthermalized : {therm}\n"
    R = len(mcs)
    dt = t - prev["t"]
    for n in range(R):
        att_local = dt * mcs[0].lattice.size / (p.overrelaxation_rate or 1)
        loc = (stats["accepted_local"][n] - prev["local"][n]) / att_local * 100.0 if att_local else 0.0
        if R == 1:
            s += f"\t\tupdate acceptance rate : {loc:.2f}%\tsigma : {stats['sigma'][n]:.2f}\n"
        else:
            att_sw = dt / p.swap_rate / (2.0 if n in (0, R - 1) else 1.0)
            ex = (stats["accepted_exchange"][n] - prev["exch"][n]) / att_sw * 100.0 if att_sw else 0.0
            s += (f"\t\tsimulation {n} update acceptance rate : {loc:.2f}%\tsigma : {stats['sigma'][n]:.2f}\n"
                  f"\t\tsimulation {n} replica exchange acceptance rate : {ex:.2f}%\n")
    print(s)
    prev["t"] = t
    prev["local"] = stats["accepted_local"].copy()
    prev["exch"] = stats["accepted_exchange"].copy()


def parallel_tempering(mcs, saveIC=(), alg=None, verbose=True, resume=False):
    """parallel_tempering! (monte_carlo.jl:235-398) for the replicas of one
    process.  `mcs` is one MonteCarlo per temperature (the reference's one
    per MPI rank; a single MonteCarlo runs without exchanges, as the
    reference does with commSize 1).  Replicas run as one device batch;
    after the call mcs[r] holds the configuration at temperature slot r and
    the slot's observables.

    With output files, every checkpoint also saves the whole run (group
    `resume` of the first replica's file; the copy to the host overlaps the
    sweeps that follow), and resume=True continues from that save: the
    result is the one an uninterrupted run gives.  The reference has no
    resume."""
    single = isinstance(mcs, MonteCarlo)
    mcs = [mcs] if single else list(mcs)
    alg = alg or Metropolis()
    kind = _kind(alg)
    if kind == _lib.CONSTRAINT_ADAPTIVE:
        raise NotImplementedError(
            "MetropolisConstraintAdaptive in parallel tempering: the device PT loop runs every sweep of a "
            "replica in one persistent kernel, and the constraint's per-proposal multiplier chain is run "
            "by its own kernels (csmc_sweep / simulated_annealing)")
    p = mcs[0].parameters
    R = len(mcs)
    if R == 1:
        warnings.warn("MPI commSize of 1; no replica exchanges will occur!")
    if p.report_interval == 0:
        raise ZeroDivisionError("report_interval = 0: the reference fails with DivideError at "
                                "`sweep % report_interval` (monte_carlo.jl:383)")
    if p.t_measurement > 0 and (p.checkpoint_rate == 0 or p.probe_rate == 0):
        raise ZeroDivisionError("checkpoint_rate and probe_rate must be > 0: the reference fails "
                                "with DivideError once sweep >= t_thermalization (monte_carlo.jl:355,368)")
    if R > 1 and p.swap_rate == 0:
        raise ZeroDivisionError("swap_rate = 0: DivideError in the reference (monte_carlo.jl:308)")
    lat = mcs[0].lattice
    for m in mcs:
        if m.lattice.size != lat.size or m.lattice.unit_cell is not lat.unit_cell and \
                m.lattice.shape != lat.shape:
            raise ValueError("all replicas must share one lattice")
    temps = np.array([m.T for m in mcs])
    b = SpinBatch([m.lattice for m in mcs], precision=mcs[0].precision, ctx=mcs[0].ctx)
    for r, m in enumerate(mcs):
        b.set_spins(r, m.lattice._spins)
    params = p.as_dict()
    b.pt_init(temps, params, kind, sigma0=mcs[0].sigma0, seed=mcs[0].seed)
    # corr = true: S(k) at every probe sweep, averaged per temperature slot
    # on the device (monte_carlo.jl:371-375)
    ks = mcs[0].momentum_vectors
    corr = mcs[0].corr and ks.shape[1] > 0
    if corr:
        b.pt_set_correlations(ks)
    total = p.t_thermalization + p.t_measurement
    out = any(m.outpath for m in mcs)
    from . import hdf5io
    if resume:
        saved = hdf5io.read_resume_state(mcs[0])
        if saved is not None:
            b.pt_load(saved[0])
            if verbose:
                print(f"Resuming at sweep {saved[1]} from {mcs[0].outpath}")
    saving = [False]

    def finish_save():
        # the save started at the last checkpoint, written once it is on the host
        if saving[0]:
            state = b.pt_save_wait()
            hdf5io.write_resume_state(mcs[0], state, saving[1])
            saving[0] = False
    if verbose:
        print(f"Running sweeps on {datetime.datetime.now():%d %b %Y %H:%M:%S}.")
    prev = {"t": 0, "local": np.zeros(R), "exch": np.zeros(R)}
    while True:
        stopped, event = b.pt_advance(-1)
        if event == 1 and (out or saveIC):
            st = b.pt_state()
            for r, m in enumerate(mcs):
                m.lattice._spins[:] = b.get_spins(int(st["system_of_slot"][r]))
                if m.outpath:
                    hdf5io.write_MC_checkpoint(m)
                if r in saveIC:
                    k = (stopped - p.t_thermalization) // p.checkpoint_rate
                    hdf5io.write_initial_configuration(hdf5io.ic_path(m, r, k), m)
            if mcs[0].outpath:
                finish_save()
                b.pt_save_async()
                saving[:] = [True, stopped]
        elif event == 2 and verbose:
            _stats_line(mcs, stopped, total, b.pt_state(), prev)
        if event == 3:
            break
    finish_save()
    st = b.pt_state()
    obs = b.pt_observables()
    if corr:
        cmean, ccount = b.pt_correlations()
    for r, m in enumerate(mcs):
        m.lattice._spins[:] = b.get_spins(int(st["system_of_slot"][r]))
        m.sigma = float(st["sigma"][r])
        m.observables = {k: float(v[r]) for k, v in obs.items()}
        m.pt_stats = {k: (v[r] if np.ndim(v) else v) for k, v in st.items()}
        if m.corr:
            # mean(mc.observables.correlations): 9 x n_k (9 x 0 without momenta)
            m.correlations = cmean[r] if corr else np.zeros((9, 0))
            m.correlation_count = ccount if corr else 0
    if out:
        for m in mcs:
            if m.outpath:
                hdf5io.write_final_observables(m)
    if verbose:
        print(f"Simulation finished on {datetime.datetime.now():%d %b %Y %H:%M:%S}.")
    b.close()
    return mcs[0] if single else mcs
