"""Benchmark of the sweep hot path on one MI355X.

Default workload: honeycomb Kitaev-Heisenberg model (J = 1, K = -2 on the
x/y/z bonds), L = 64 (8192 spins), S = 1/2, parallel tempering with 64
replicas at geometric temperatures 0.01 .. 1.0, exchange every 10 sweeps,
one Metropolis sweep every 10 overrelaxation sweeps, a probe of E and |M|
every 10 sweeps after thermalisation; FP32 spin updates, FP64 energies,
exchange decisions and observable binning.

One step = one exchange period: 10 overrelaxation sweeps + 1 Metropolis
sweep + FP64 total energy of every replica + the replica exchange (+ the
probe), i.e. 11 * 64 * 8192 = 5,767,168 spin updates.  The 1e4
thermalisation sweeps run first, untimed; then 1024 untimed warm-up steps
(one chained launch of the sweep kernel) and the timed steps, all in the
measurement phase, so every timed step carries its probe.  The default
timed run is the configured 1e5 measurement sweeps (10,000 steps).
Metric: spin-updates/s.

Beside the batched FP32 form it measures the plain FP64 GPU form (one
replica at a time, one launch per colour pass) on the same lattice, and the
oracle (CPU FP64 restatement of the reference's PT loop: random-site
Metropolis, sequential overrelaxation, total energy per Metropolis sweep,
even/odd exchange) on a bounded sample on the host cores.

Other workloads (--workload ...), each printing one line with the same
fields:
  anneal_square   the reference's CPU-runnable case: square L=16, simulated
                  annealing to 1e-3 and deterministic updates, FP64
  anneal_uneven   256 independent lattices of uneven size annealed in one
                  launch
  pt_pyrochlore   the pyrochlore model with cubic, quartic and onsite terms
                  under parallel tempering
  adjoint         projected gradient descent on 128 triangular J1-J2
                  lattices, L=96, FP64
  ssf             the equal-time structure factor of the 64 replicas of the
                  default workload at the 65 x 65 allowed momenta of one
                  zone copy, against the FP64 MFMA roofline
  probe_L512      the default geometry at L=512 (403 MB of spin records,
                  beyond the Infinity Cache) as batch-wide tiled sweeps (two
                  overrelaxation sweeps a launch)
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "classicalspinmc.jl_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "spin-updates/s on 1 MI355X (batched FP32 PT vs plain); % of HBM roofline"
PEAK_HBM_GBS = 8000.0
BYTES_PER_UPDATE_FP32 = 24  # read + write own spin, 3 x 4 B each
# MI355X compute figures for the issue-rate fractions: 256 CUs at a 2.4 GHz
# peak engine clock; a CU issues 64 non-packed FP32 VALU lane-operations per
# clock (4 SIMDs x 16 lanes; packed FP32, which would double it, is not used
# by this build) and its LDS serves 128 B per clock
CHIP_CUS = 256
CLOCK_HZ = 2.4e9
# The 157.3 TFLOP/s FP32 figure (128 FMA lanes per CU and clock) needs
# packed FP32; non-packed VALU code issues a wave64 instruction every four
# cycles per SIMD.  The tiled sweep kernel meets that ceiling: its counters
# show 39.7 x 10^12 lane operations per second (1.01 x the four-cycle rate
# at 2.4 GHz), and a build with 26 % fewer VALU instructions ran 15 % faster.
VALU_LANES_PER_CU_CLK = 64
LDS_BYTES_PER_CU_CLK = 128
LDS_BYTES_PER_LANE_INST = 16  # the register-table kernel's LDS traffic is 16-byte spin records


def issue_fractions(sq, updates_per_dispatch, value, cus_busy, lds_bytes_per_lane_inst=LDS_BYTES_PER_LANE_INST):
    """VALU and LDS fractions of the register-table kernel from its counters
    (SQ_INSTS_VALU, SQ_INSTS_LDS per dispatch, wave instructions, 64 lanes
    each) at the live throughput `value` (updates/s): against the whole chip
    and against the CUs the kernel occupies."""
    valu_per_update = sq["SQ_INSTS_VALU"] * 64 / updates_per_dispatch
    lds_per_update = sq["SQ_INSTS_LDS"] * 64 * lds_bytes_per_lane_inst / updates_per_dispatch
    valu = valu_per_update * value
    lds = lds_per_update * value
    vpeak = VALU_LANES_PER_CU_CLK * CLOCK_HZ
    lpeak = LDS_BYTES_PER_CU_CLK * CLOCK_HZ
    return {
        "valu": {"lane_ops_per_update": round(valu_per_update, 2), "achieved": valu / 1e12,
                 "unit": "T lane-ops/s", "peak": CHIP_CUS * vpeak / 1e12,
                 "frac": round(valu / (CHIP_CUS * vpeak), 4),
                 "frac_busy_cus": round(valu / (cus_busy * vpeak), 4), "busy_cus": cus_busy},
        "lds": {"bytes_per_update": round(lds_per_update, 1), "achieved": lds / 1e12, "unit": "TB/s",
                "peak": CHIP_CUS * lpeak / 1e12, "frac": round(lds / (CHIP_CUS * lpeak), 4),
                "frac_busy_cus": round(lds / (cus_busy * lpeak), 4), "busy_cus": cus_busy},
        "issue_note": "instruction counts per update from the kernel's SQ counters (profiles/), times "
                      "the live updates/s; the kernel keeps the spins in LDS, so VALU issue and LDS "
                      "bandwidth on the CUs it occupies bound it, not HBM"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_probe(nbytes, reps=5):
    """device-to-device copies of a known size, to calibrate the FETCH_SIZE /
    WRITE_SIZE counters against bytes actually moved"""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def baseline_oracle():
    """The oracle as the CPU baseline: its speed build (liboracle_bench.so,
    -O3 for x86-64-v3), never the product path"""
    import oracle.oracle as orc
    if orc._lib is None:
        orc.use_bench_build()
    return orc


def cpu_baseline(uc, L, S, temps, seconds, threads):
    """The oracle's reference-algorithm PT loop (orc_pt_reference_run) on the
    host: the replicas split into `threads` blocks, one host thread per block
    (the oracle's C loop runs without the GIL), each block exchanging within
    itself.  Bounded to about `seconds` of wall time; returns updates/s and a
    description of the sample."""
    orc = baseline_oracle()
    T = orc.Tables(uc, (L, L))
    R = len(temps)
    threads = max(1, min(threads, R))
    blocks = np.array_split(np.arange(R), threads)
    counts = [0] * threads
    sweeps = [0] * threads
    stop = time.perf_counter() + seconds

    def work(k):
        idx = blocks[k]
        v = np.random.default_rng(k).normal(size=(len(idx), T.N, 3))
        spins = np.ascontiguousarray(S * v / np.linalg.norm(v, axis=2, keepdims=True))
        E = np.zeros(len(idx))
        perm = np.arange(len(idx), dtype=np.intc)
        tb = np.ascontiguousarray(temps[idx])
        sw = 0
        while time.perf_counter() < stop and sw < 100000:
            counts[k] += orc.pt_reference_run(T, spins, tb, S, sw, 10, 10, 10, 2 + k, E, perm)
            sw += 10
        sweeps[k] = sw

    t0 = time.perf_counter()
    pool = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    dt = time.perf_counter() - t0
    total = sum(counts)
    sample = (f"oracle PT loop (random-site Metropolis + sequential overrelaxation + total_energy "
              f"per Metropolis sweep + even/odd exchange; C, gcc -O3 -march=x86-64-v3), {R} replicas "
              f"of the same lattice in {threads} blocks of {R // threads}-{-(-R // threads)} replicas, "
              f"one host thread per block, exchanges within a block, no message cost; "
              f"{min(sweeps)}-{max(sweeps)} sweeps per block in {dt:.1f} s")
    return total / dt, sample


def host_info():
    """CPU model and core counts of the host the baseline ran on"""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def finish(out, args):
    print(json.dumps(out))


def fp32_error(csmc, W, lat, temps, ctx, t_th=2000, t_meas=20000):
    """Error of the FP32 path against FP64 on the headline lattice and ladder:
    - evaluation_energy_rel: the FP64 total energy the FP32 batch reports for
      its own state against an FP64 batch holding the same spins (the
      reduction is FP64, so this is the error of an FP32 energy result);
    - per slot, <E> of an FP32 and an FP64 parallel-tempering run of the same
      schedule (t_th thermalisation + t_meas measurement sweeps, independent
      seeds) with their binning errors: the largest relative difference
      |<E>_32 - <E>_64| / |<E>_64|, the combined error bar at that slot, and
      the largest |z| = |difference| / combined error over the slots."""
    R = len(temps)
    p = dict(t_thermalization=t_th, t_measurement=t_meas, probe_rate=10, swap_rate=10, overrelaxation_rate=10,
             report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    runs = {}
    for prec, seed in ((32, 41), (64, 43)):
        b = csmc.SpinBatch([lat] * R, precision=prec, ctx=ctx)
        b.randomize(seed)
        b.pt_init(temps, p, csmc.METROPOLIS, seed=seed)
        while b.pt_advance(-1)[1] != 3:
            pass
        runs[prec] = (b, b.pt_observables())
    b32, o32 = runs[32]
    b64, o64 = runs[64]
    diff = o32["energy"] - o64["energy"]
    err = np.sqrt(o32["energy_err"] ** 2 + o64["energy_err"] ** 2)
    rel = np.abs(diff) / np.abs(o64["energy"])
    k = int(np.argmax(rel))
    out = {"per_slot_energy_rel_max": float(rel[k]), "at_slot": k,
           "combined_error_rel_at_slot": float(err[k] / abs(o64["energy"][k])),
           "max_abs_z": float(np.max(np.abs(diff) / err)),
           "slots_beyond_3_sigma": int(np.sum(np.abs(diff) > 3 * err)), "slots": int(R),
           "sweeps": f"{t_th} + {t_meas} per run, FP32 and FP64 PT with independent seeds"}
    # the FP32 batch's own energy against FP64 evaluation of the same spins
    for i in range(R):
        b64.set_spins(i, b32.get_spins(i))
    e32, e64 = b32.total_energy(), b64.total_energy()
    out["evaluation_energy_rel"] = float(np.max(np.abs(e32 - e64) / np.abs(e64)))
    b32.close()
    b64.close()
    return out


def checkpoint_cost(csmc, b, temps):
    """Cost of saving the whole PT state of the batch: a synchronous save,
    and an asynchronous one queued ahead of the next chained launch (1024
    exchange periods); the host time it exposes is the start call plus the
    wait after that launch, and the launch's own time shows the copy's share
    of the stream."""
    # t_thermalization = 1: with 0, sweep 0 would be a checkpoint sweep
    # (sweep >= t_thermalization and sweep % checkpoint_rate == 0) and the
    # first advance would stop there
    params = dict(t_thermalization=1, t_measurement=3 * 10240, probe_rate=10, swap_rate=10,
                  overrelaxation_rate=10, report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    b.pt_init(temps, params, csmc.METROPOLIS, seed=3)
    assert b.pt_advance(10240)[0] == 10240
    # first use allocates the page-locked buffer: not timed
    b.pt_save()
    b.pt_save_async()
    b.pt_save_wait()
    t0 = time.perf_counter()
    state = b.pt_save()
    t_sync = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.pt_save_async()
    t_start = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.pt_advance(2 * 10240)
    b.sync()
    t_with = time.perf_counter() - t0
    t0 = time.perf_counter()
    same = b.pt_save_wait()
    t_wait = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.pt_advance(3 * 10240)
    b.sync()
    t_without = time.perf_counter() - t0
    assert len(same) == len(state)
    return {"state_bytes": len(state), "sync_save_ms": round(t_sync * 1e3, 3),
            "async_exposed_ms": round((t_start + t_wait) * 1e3, 3),
            "launch_ms_with_copy": round(t_with * 1e3, 3), "launch_ms_without": round(t_without * 1e3, 3),
            "note": "csmc_pt_save vs csmc_pt_save_async + csmc_pt_save_wait around one chained launch of "
                    "1024 exchange periods"}


def run_pt(args, W, csmc, uc, L, S, R, temps, precision, name, t_th=10 ** 4):
    """The PT schedule of the workload: the t_th thermalisation sweeps run
    untimed, then `warmup` untimed and `steps` timed exchange periods, all
    in the measurement phase (every period has its probe)."""
    lat = csmc.Lattice((L, L) if uc.D == 2 else (L, L, L), uc, S, seed=2)
    N = lat.size
    sweeps_per_step = 10
    meas_sweeps = (args.warmup + args.steps) * sweeps_per_step
    params = dict(t_thermalization=t_th, t_measurement=max(10 ** 5, meas_sweeps), probe_rate=10,
                  swap_rate=10, overrelaxation_rate=10, report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    ctx = csmc.default_context()
    b = csmc.SpinBatch([lat] * R, precision=precision, ctx=ctx)
    b.randomize(2)
    b.pt_init(temps, params, csmc.METROPOLIS, seed=2)
    mode = b.storage_mode()
    parts = b.pt_split_parts()
    log(f"[bench] {name}: lattice N={N} replicas={R} colours={lat.info.n_colours} storage mode {mode}, "
        f"{parts} workgroup(s) per replica")
    t0 = time.perf_counter()
    assert b.pt_advance(t_th)[0] == t_th
    ctx.sync()
    log(f"[bench] {t_th} thermalisation sweeps done (untimed, {time.perf_counter() - t0:.3f} s)")
    b.pt_advance(t_th + args.warmup * sweeps_per_step)
    ctx.sync()
    log(f"[bench] warmup {args.warmup} steps done")
    ctx.timing(True)
    start_sweep = t_th + args.warmup * sweeps_per_step
    ctx.sync()
    t0 = time.perf_counter()
    stopped, _ = b.pt_advance(start_sweep + args.steps * sweeps_per_step)
    ctx.sync()
    dt = time.perf_counter() - t0
    kern_ms, launches = ctx.timing_read()
    ctx.timing(False)
    assert stopped == start_sweep + args.steps * sweeps_per_step, stopped
    updates_per_step = (sweeps_per_step + 1) * R * N
    value = updates_per_step * args.steps / dt
    log(f"[bench] timed {args.steps} steps in {dt:.3f} s: {value / 1e9:.2f} G updates/s; "
        f"{launches} launches, {kern_ms / max(launches, 1):.3f} ms avg, "
        f"{kern_ms / args.steps * 1e3:.2f} us kernel time per step")
    st = b.pt_state()
    assert np.all(np.isfinite(st["E"])) and st["accepted_exchange"].sum() > 0
    probes = b.pt_observables()
    assert np.all(np.isfinite(probes["energy"])), "every timed period is in the measurement phase"
    s0 = b.get_spins(0)
    assert np.allclose(np.linalg.norm(s0, axis=1), S, atol=1e-4 * S)
    return dict(b=b, lat=lat, N=N, dt=dt, kern_ms=kern_ms, launches=launches, value=value,
                updates_per_step=updates_per_step, mode=mode, parts=parts, ctx=ctx)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run device-to-device copies of this many bytes (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000,
                    help="timed exchange periods (10000 = the 1e5 measurement sweeps)")
    ap.add_argument("--warmup", type=int, default=1024,
                    help="untimed steps first (1024 = one whole chained launch)")
    ap.add_argument("--workload", default="pt_honeycomb",
                    choices=["pt_honeycomb", "anneal_uneven", "pt_pyrochlore", "ssf", "probe_L512", "adjoint",
                             "anneal_square"])
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU form")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, len(os.sched_getaffinity(0))),
                    help="host threads of the CPU baseline (at most 16: the GPU box's host share "
                         "for one GPU is 16 cores)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; --gpus N > 1 is not "
                                   "supported: a multi-GPU run would shard the replica batch by process, one "
                                   "process per GPU, which this build does not do"}))
        sys.exit(2)

    import csmc
    from csmc import workloads as W

    if args.workload == "anneal_uneven":
        return bench_anneal(args, csmc, W)
    if args.workload == "pt_pyrochlore":
        return bench_pyrochlore(args, csmc, W)
    if args.workload == "ssf":
        return bench_ssf(args, csmc, W)
    if args.workload == "probe_L512":
        return bench_probe(args, csmc, W)
    if args.workload == "adjoint":
        return bench_adjoint(args, csmc, W)
    if args.workload == "anneal_square":
        return bench_anneal_square(args, csmc, W)

    L, R = args.L, args.replicas
    uc = W.honeycomb_kitaev(1.0, -2.0)
    temps = W.geometric_temperatures(0.01, 1.0, R)
    r = run_pt(args, W, csmc, uc, L, 0.5, R, temps, 32, "pt_honeycomb")
    N, ctx = r["N"], r["ctx"]
    value, kern_ms, launches = r["value"], r["kern_ms"], r["launches"]
    updates_per_step = r["updates_per_step"]

    # roofline of the dominant kernel (the PT sweep kernel; one launch runs a
    # chain of exchange periods): algorithmic bytes of a launch over its
    # average duration
    steps_per_launch = args.steps / launches
    bytes_per_launch = BYTES_PER_UPDATE_FP32 * updates_per_step * steps_per_launch
    avg_ms = kern_ms / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    fractions = None
    split = r["parts"] > 1
    kname = "k_wg_split" if split else "k_wg_reg"
    tpath = os.path.join(ROOT, "profiles", f"{kname}_pt_honeycomb_L64_R64_traffic.json")
    if L == 64 and R == 64 and os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        traffic = prof["traffic_bytes_per_step"] * steps_per_launch
        if "sq_per_dispatch" in prof:
            fractions = issue_fractions(prof["sq_per_dispatch"],
                                        updates_per_step * prof["steps_per_dispatch"], value,
                                        min(CHIP_CUS, R * r["parts"]))
    kdesc = (f"k_wg_split (register-table PT sweep kernel, each replica split over {r['parts']} "
             "workgroups with halos, chained exchange periods)" if split else
             "k_wg_reg (register-table PT sweep kernel, one workgroup per replica, chained "
             "exchange periods)")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "kernel": kdesc,
                "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_ms, 4),
                "steps_per_launch": round(steps_per_launch, 2),
                "note": "spins stay in LDS for a whole launch (a chain of up to 1024 exchange "
                        "periods): HBM traffic is far below the algorithmic 24 B per update"}
    if fractions:
        roofline.update(fractions)

    out = {
        "metric": METRIC,
        "value": value, "unit": "spin-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["dt"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"honeycomb Kitaev-Heisenberg J=1 K=-2, L={L} ({N} spins), S=1/2, "
                               f"parallel tempering, {R} replicas, T=0.01..1 geometric, "
                               "swap_rate=10, overrelaxation_rate=10, probe_rate=10",
                   "replicas": R, "sites": N, "updates_per_step": updates_per_step,
                   "precision": "fp32 spins, fp64 energies/exchange/binning",
                   "storage_mode": r["mode"], "workgroups_per_replica": r["parts"]},
        "roofline": roofline,
    }
    out["checkpoint"] = checkpoint_cost(csmc, r["b"], temps)
    log(f"[bench] checkpoint: {out['checkpoint']}")
    r["b"].close()
    out["precision"] = fp32_error(csmc, W, r["lat"], temps, ctx)

    if not args.no_plain:
        # plain unbatched FP64 form: one replica at a time, one launch per
        # colour pass; one exchange period (10 OR + 1 Metropolis) on 4 replicas
        lat = r["lat"]
        pb = csmc.SpinBatch([lat] * 4, precision=64, ctx=ctx)
        pb.randomize(3)
        Tp = temps[:4]
        pb.sweep(csmc.OVERRELAX, n_sweeps=2, plain=True)
        ctx.sync()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            pb.sweep(csmc.OVERRELAX, n_sweeps=10, plain=True)
            pb.sweep(csmc.METROPOLIS, T=Tp, n_sweeps=1, seed=5, plain=True)
            pb.total_energy()
        ctx.sync()
        dtp = time.perf_counter() - t0
        plain = reps * 11 * 4 * N / dtp
        # the same period with the sweeps captured as hipGraphs (one per
        # kind) and replayed
        g_or = pb.graph_sweeps(csmc.OVERRELAX, n_sweeps=10)
        g_me = pb.graph_sweeps(csmc.METROPOLIS, T=Tp, n_sweeps=1, seed=5)
        g_or.launch()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            g_or.launch()
            g_me.launch()
            pb.total_energy()
        ctx.sync()
        dtg = time.perf_counter() - t0
        graphed = reps * 11 * 4 * N / dtg
        g_or.close()
        g_me.close()
        out["plain_gpu"] = {"value": plain, "unit": "spin-updates/s", "dtype": "fp64",
                            "form": "one replica at a time, one launch per colour pass",
                            "batched_over_plain": value / plain,
                            "graph_captured": graphed}
        pb.close()
        log(f"[bench] plain FP64 form: {plain / 1e9:.3f} G updates/s, captured as graphs {graphed / 1e9:.3f}")

    if not args.no_cpu:
        cpu, sample = cpu_baseline(W.honeycomb_kitaev(1.0, -2.0), L, 0.5, temps, args.cpu_seconds,
                                   args.cpu_threads)
        out["cpu_baseline"] = {"value": cpu, "unit": "spin-updates/s",
                               "cores": max(1, min(args.cpu_threads, R)), "kind": "port",
                               "sample": sample, "gpu_over_cpu": value / cpu, **host_info()}
        log(f"[bench] oracle CPU: {cpu / 1e6:.1f} M updates/s on {out['cpu_baseline']['cores']} threads")

    finish(out, args)


def bench_anneal(args, csmc, W):
    """256 independent square / triangular lattices (L in 16..128, per-lattice
    DM couplings and fields) annealed 0.9^x from 1 to 0.05 (29 temperatures),
    690 sweeps per temperature (about 2e4 sweeps per lattice), FP32, all in
    one batch; timed over the whole schedule."""
    cells = W.uneven_batch(n=256, seed=3)
    lats = [csmc.Lattice(shape, uc, 1.0, seed=3 + k) for k, (uc, shape) in enumerate(cells)]
    temps = W.anneal_temperatures(1.0, 0.9, 0.05)
    t_th, or_rate = 690, 10
    ctx = csmc.default_context()
    b = csmc.SpinBatch(lats, precision=32, ctx=ctx)
    b.randomize(3)
    n_total = sum(l.size for l in lats)
    log(f"[bench] anneal_uneven: {len(lats)} lattices, {n_total} spins, storage mode {b.storage_mode()}")
    b.anneal(temps[:1], 20, or_rate, csmc.METROPOLIS, seed=1)  # warm-up launch
    ctx.sync()
    ctx.timing(True)
    t0 = time.perf_counter()
    b.anneal(temps, t_th, or_rate, csmc.METROPOLIS, seed=3)
    ctx.sync()
    dt = time.perf_counter() - t0
    kern_ms, launches = ctx.timing_read()
    per_T = (t_th - 1) + (t_th - 1) // or_rate
    updates = per_T * len(temps) * n_total
    value = updates / dt
    achieved = BYTES_PER_UPDATE_FP32 * updates / (kern_ms * 1e-3) / 1e9
    out = {"metric": METRIC, "value": value, "unit": "spin-updates/s", "n_gpus": 1, "steps": len(temps),
           "warmup": 1, "ms_per_step": dt / len(temps) * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": "256 independent lattices (128 square + 128 triangular, L in 16..128, "
                                  "DM couplings and fields per lattice) annealed 0.9^x to 0.05, "
                                  "690 sweeps per temperature, overrelaxation_rate=10",
                      "sites": n_total, "temperatures": len(temps)},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                        "kernel": "k_wg (whole anneal in one launch)"}}
    log(f"[bench] {value / 1e9:.2f} G updates/s over {dt:.2f} s")
    finish(out, args)


def bench_pyrochlore(args, csmc, W):
    """Pyrochlore L=16 (16384 spins): NN Heisenberg, biquadratic quartic,
    tetrahedron-face chirality (cubic) and <111> onsite anisotropy; 32
    replicas at geometric T 0.02..2, FP32 PT with the config-2 schedule."""
    R = 32
    temps = W.geometric_temperatures(0.02, 2.0, R)
    steps = min(args.steps, 500)
    args.steps = steps
    args.warmup = min(args.warmup, 20)
    r = run_pt(args, W, csmc, W.pyrochlore_mixed(), 16, 1.0, R, temps, 32, "pt_pyrochlore")
    value = r["value"]
    # roofline of the PT kernel: algorithmic bytes of a launch over its
    # average duration, beside its VALU issue (the counters per exchange
    # period come from profiles/)
    ups = r["updates_per_step"]
    steps_per_launch = steps / r["launches"]
    avg_ms = r["kern_ms"] / r["launches"]
    achieved = BYTES_PER_UPDATE_FP32 * ups * steps_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                "kernel": "k_wg (storage mode 0, chained PT, parts of a replica on one XCD)",
                "avg_launch_ms": round(avg_ms, 4), "steps_per_launch": steps_per_launch}
    tpath = os.path.join(ROOT, "profiles", "k_wg_pt_pyrochlore_L16_R32_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        roofline["traffic"] = prof["traffic_bytes_per_step"] * steps_per_launch
        parts = max(1, r.get("parts", 1))
        roofline.update(issue_fractions(prof["sq_per_dispatch"], ups * prof["steps_per_dispatch"], value,
                                        min(CHIP_CUS, R * parts), lds_bytes_per_lane_inst=4))
        roofline["issue_note"] = ("instruction counts per update from the kernel's SQ counters (profiles/), "
                                  "times the live updates/s; the spins stay in device memory (L2), so the "
                                  "gathers' latency and VALU issue on the busy CUs bound it")
    out = {"metric": METRIC, "value": value, "unit": "spin-updates/s", "n_gpus": 1, "steps": steps,
           "warmup": args.warmup, "ms_per_step": r["dt"] / steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": "pyrochlore L=16 (16384 spins), J1 + biquadratic + chirality + onsite, "
                                  "parallel tempering, 32 replicas, T=0.02..2",
                      "storage_mode": r["mode"], "parts_per_replica": r.get("parts", 1)},
           "roofline": roofline}
    r["b"].close()
    finish(out, args)


def bench_probe(args, csmc, W):
    """Roofline probe: the headline geometry at L=512 (524,288 sites) x 64
    FP32 replicas, 403 MB of packed spin records -- more than the 256 MB
    Infinity Cache, so every sweep streams the state from HBM.  One step = 10
    overrelaxation sweeps + 1 Metropolis sweep of every replica in the
    batch-wide form (csmc_sweep_passes: one launch per sweep of LDS tiles with
    recomputed halos, partners by index arithmetic); the persistent kernel
    (storage mode 0, packed schedule) on the same state is timed beside it."""
    L, R = 512, 64
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice((L, L), uc, 0.5, seed=2)
    N = lat.size
    temps = W.geometric_temperatures(0.01, 1.0, R)
    ctx = csmc.default_context()
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(2)
    steps = max(1, min(args.steps, 20))
    b.sweep_passes(csmc.OVERRELAX, n_sweeps=2)
    ctx.sync()
    ctx.timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.sweep_passes(csmc.OVERRELAX, n_sweeps=10)
        b.sweep_passes(csmc.METROPOLIS, T=temps, n_sweeps=1, seed=3)
    ctx.sync()
    dt = time.perf_counter() - t0
    kern_ms, launches = ctx.timing_read()
    ctx.timing(False)
    updates = steps * 11 * R * N
    value = updates / dt
    achieved = BYTES_PER_UPDATE_FP32 * updates / (kern_ms * 1e-3) / 1e9
    tiles, why = b.tiles()
    traffic, issue = None, {}
    tpath = os.path.join(ROOT, "profiles", "k_tile_probe_honeycomb_L512_R64_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        # HBM bytes per launch of the overrelaxation kernel (two sweeps of
        # every replica from one staging of each tile)
        traffic = tj["traffic_bytes_per_update"] * tj["updates_per_dispatch"]
        # the overrelaxation kernel's VALU issue at its own live rate (10 of
        # the 11 sweeps of a step)
        ms_or = tile_or_ms(csmc, b, ctx)
        # LDS: a packed 12-byte record moves in two instructions (8 + 4 B)
        issue = issue_fractions(tj["sq_per_dispatch"], tj["updates_per_dispatch"], R * N / (ms_or * 1e-3),
                                CHIP_CUS, lds_bytes_per_lane_inst=6)
        issue["issue_note"] = ("instruction counts of the overrelaxation tile kernel from its SQ counters "
                               "(profiles/), times its live updates/s (%.1f us per sweep): VALU issue is "
                               "the kernel's main limit (SALU, LDS latency and the barriers between colour "
                               "steps the rest); two sweeps a launch move %.1f measured HBM bytes per "
                               "update, below the 24-byte algorithmic figure"
                               % (ms_or * 1e3, tj["traffic_bytes_per_update"]))
    # the persistent kernel on the same state (storage mode 0: the packed
    # schedule splits each replica into parts on one XCD)
    b.sweep(csmc.OVERRELAX, n_sweeps=1)
    ctx.sync()
    t0 = time.perf_counter()
    b.sweep(csmc.OVERRELAX, n_sweeps=4)
    ctx.sync()
    persistent = 4 * R * N / (time.perf_counter() - t0)
    # the autotuner on the same state (it restores the state)
    tune_times, tune_choice = b.autotune(csmc.OVERRELAX, n_sweeps=2)
    out = {"metric": METRIC, "value": value, "unit": "spin-updates/s", "n_gpus": 1, "steps": steps, "warmup": 1,
           "ms_per_step": dt / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"roofline probe: honeycomb Kitaev-Heisenberg J=1 K=-2, L={L} ({N} spins), "
                                  f"{R} FP32 replicas ({R * (N + 1) * 12 / 1e6:.0f} MB of spin records), 10 "
                                  "overrelaxation + 1 Metropolis sweep per step, batch-wide tiled sweeps",
                      "sites": N, "replicas": R, "storage_mode": b.storage_mode(),
                      "tiles_per_replica_one_sweep_plan": tiles, "sweeps_per_tiled_launch": b.tile_sweeps()},
           "roofline": dict({"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                             "kernel": "k_tile<float> over all replicas (one launch per two overrelaxation "
                                       "sweeps, one per Metropolis sweep)" if tiles else
                                       f"k_pass<float> over all replicas (one launch per colour pass: {why})",
                             "note": "achieved = 24 algorithmic bytes per update x updates / kernel time; "
                                     "traffic = measured HBM bytes per launch of the overrelaxation tile "
                                     "kernel (two sweeps of all replicas; FETCH_SIZE + WRITE_SIZE, "
                                     "copy-kernel scaled)"},
                            **issue),
           "persistent_kernel": {"value": persistent, "unit": "spin-updates/s",
                                 "form": "persistent kernel, storage mode 0, packed schedule (parts of a replica "
                                         "on one XCD, spins gathered from device memory), overrelaxation sweeps"},
           "autotune": {"ms_for_2_overrelaxation_sweeps": {k: round(v, 3) for k, v in tune_times.items()},
                        "chosen": tune_choice}}
    b.close()
    log(f"[bench] probe: {value / 1e9:.2f} G updates/s, {achieved:.0f} GB/s algorithmic; persistent "
        f"{persistent / 1e9:.2f} G updates/s")
    finish(out, args)


def tile_or_ms(csmc, b, ctx, sweeps=20):
    """average ms of one overrelaxation sweep of the batch-wide form, timed
    with the library's events around the launches"""
    b.sweep_passes(csmc.OVERRELAX, n_sweeps=2)
    ctx.sync()
    ctx.timing(True)
    b.sweep_passes(csmc.OVERRELAX, n_sweeps=sweeps)
    ms, _ = ctx.timing_read()
    ctx.timing(False)
    return ms / sweeps


def bench_anneal_square(args, csmc, W):
    """The reference's CPU-runnable case: square lattice L=16, S=1, J=-I on
    the 4 bonds, h=0.1 z, FP64; simulated_annealing with T=0.9^x from 1.0
    down to 1e-3 (66 temperatures), 1e4 thermalization sweeps per
    temperature with a Metropolis sweep every 10, then 1e4 deterministic
    updates.  Timed end to end through the MonteCarlo driver; the ground
    state E/N = -2S^2|J| - hS = -2.1 is checked."""
    import contextlib
    orc = baseline_oracle()
    uc = W.square_heisenberg(-1.0, (0.0, 0.0, 0.1))
    lat = csmc.Lattice((16, 16), uc, 1.0, seed=1)
    N = lat.size
    params = {"t_thermalization": 10 ** 4, "overrelaxation_rate": 10, "t_deterministic": 10 ** 4}
    ctx = csmc.default_context()
    mc = csmc.MonteCarlo(1e-3, lat, dict(params), precision=64, seed=1, ctx=ctx)
    sched = lambda x: 0.9 ** x  # noqa: E731
    temps = []
    T = 1.0
    while T > 1e-3:
        temps.append(T)
        T = sched(len(temps))
    ctx.timing(True)
    t0 = time.perf_counter()
    with contextlib.redirect_stdout(sys.stderr):  # the driver prints one line per temperature
        csmc.simulated_annealing(mc, sched, 1.0)
        csmc.deterministic_updates(mc)
    dt = time.perf_counter() - t0
    kern_ms, launches = ctx.timing_read()
    ctx.timing(False)
    per_T = (params["t_thermalization"] - 1) + (params["t_thermalization"] - 1) // params["overrelaxation_rate"]
    updates = per_T * len(temps) * N
    value = updates / dt
    e = csmc.energy_density(mc.lattice)
    achieved = 48.0 * updates / (kern_ms * 1e-3) / 1e9
    out = {"metric": METRIC, "value": value, "unit": "spin-updates/s", "n_gpus": 1, "steps": len(temps),
           "warmup": 0, "ms_per_step": dt / len(temps) * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
           "config": {"workload": "square L=16, S=1, J=-I, h=0.1 z; simulated annealing 0.9^x from 1 to 1e-3 "
                                  "(66 temperatures x 1e4 sweeps, Metropolis every 10), then 1e4 "
                                  "deterministic updates; FP64",
                      "sites": N, "temperatures": len(temps), "final_energy_per_site": e},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": None,
                        "kernel": "k_wg<double> (one workgroup, one launch per temperature)",
                        "note": "256 sites on one CU: the sweep is latency-bound, not HBM-bound"}}
    if not args.no_cpu:
        # the oracle's reference-order loop (random-site Metropolis every 10th
        # sweep, sequential overrelaxation) at the first temperature, bounded
        T = orc.Tables(uc, (16, 16))
        spins = np.ascontiguousarray(lat._spins.copy()[None, :, :])
        E = np.zeros(1)
        perm = np.zeros(1, dtype=np.intc)
        count, sw = 0, 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_seconds, 8.0):
            count += orc.pt_reference_run(T, spins, [1.0], 1.0, sw, 100, 10, 10 ** 9, 7, E, perm)
            sw += 100
        cpu = count / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": cpu, "unit": "spin-updates/s", "cores": 1, "kind": "port",
                               "sample": f"oracle reference loop at T=1, {sw} sweeps of 256 sites",
                               "gpu_over_cpu": value / cpu}
    log(f"[bench] anneal_square: {value / 1e6:.1f} M updates/s, E/N = {e:.5f} (ground state -2.1)")
    finish(out, args)


def bench_adjoint(args, csmc, W):
    """Differentiation workload: triangular J1-J2 lattices, L=96 (9216 sites),
    S=1, J1=1 and J2 ~ U(0, 0.2) per configuration, a batch of 128 seeded FP64
    configurations.  One step = one projected gradient-descent step of every
    configuration on the device: k_gd_step computes dE/dS from the old spins
    and writes the updated spins to a second buffer.  Roofline: 48
    algorithmic bytes per site (read the FP64 spin, write the FP64 spin)."""
    orc = baseline_oracle()
    rng = np.random.default_rng(5)
    n, L = 128, 96
    J2 = rng.uniform(0.0, 0.2, size=n)
    cells = [W.triangular_j1j2(1.0, float(j)) for j in J2]
    lats = [csmc.Lattice((L, L), uc, 1.0, seed=5 + k) for k, uc in enumerate(cells)]
    N = lats[0].size
    ctx = csmc.default_context()
    b = csmc.SpinBatch(lats, precision=64, ctx=ctx)
    b.randomize(5)
    steps = max(10, min(args.steps, 200))
    b.gradient_descent(0.05, 2)
    ctx.sync()
    ctx.timing(True)
    t0 = time.perf_counter()
    b.gradient_descent(0.05, steps)
    dt = time.perf_counter() - t0
    kern_ms, launches = ctx.timing_read()
    ctx.timing(False)
    value = steps * n * N / dt
    avg_ms = kern_ms / launches
    achieved = 48.0 * n * N / (avg_ms * 1e-3) / 1e9
    traffic = None  # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled)
    tpath = os.path.join(ROOT, "profiles", "k_gd_step_adjoint_tri_L96_B128_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
    t0 = time.perf_counter()
    for _ in range(5):
        b.grad_couplings()
    dj_ms = (time.perf_counter() - t0) / 5 * 1e3
    out = {"metric": METRIC, "value": value, "unit": "spin-updates/s", "n_gpus": 1, "steps": steps, "warmup": 2,
           "ms_per_step": dt / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp64", "data": "synthetic",
           "config": {"workload": f"projected gradient descent on the sphere, {n} triangular J1-J2 lattices "
                                  f"L={L} ({N} sites), J1=1, J2~U(0,0.2), FP64, dE/dS on the device each step",
                      "systems": n, "sites": N},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                        "kernel": "k_gd_step<double> (dE/dS and the projected update of every site)",
                        "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": 48 * n * N},
           "dEdJ_ms": round(dj_ms, 3)}
    if not args.no_cpu:
        # the oracle's local field (= dE/dS for this model) of every site,
        # configurations spread over host threads (the C loop runs without
        # the GIL), bounded in time
        threads = max(1, args.cpu_threads)
        tabs = [orc.Tables(cells[k], (L, L)) for k in range(threads)]
        spins = [b.get_spins(k) for k in range(threads)]
        calls = [0] * threads
        stop = time.perf_counter() + min(args.cpu_seconds, 8.0)

        def work(k):
            while time.perf_counter() < stop:
                orc.local_fields(tabs[k], spins[k])
                calls[k] += 1

        t0 = time.perf_counter()
        pool = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        for t in pool:
            t.start()
        for t in pool:
            t.join()
        cpu = sum(calls) * N / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": cpu, "unit": "site gradients/s", "cores": threads, "kind": "port",
                               "sample": f"oracle local_fields (dE/dS) of whole configurations, {threads} host "
                                         f"threads, {sum(calls)} configurations",
                               "gpu_over_cpu": (n * N / (avg_ms * 1e-3)) / cpu}
    b.close()
    out["thermal_dE_dJ2"] = thermal_dE_dJ2(csmc, W)
    log(f"[bench] adjoint: {value / 1e9:.2f} G spin-updates/s, k_gd_step {avg_ms:.3f} ms ({achieved:.0f} GB/s); "
        f"thermal d<E>/dJ2 {out['thermal_dE_dJ2']}")
    finish(out, args)


def thermal_dE_dJ2(csmc, W, L=96, n=128, T=0.1, J2=0.1, d=0.005, samples=80):
    """The thermal gradient of the differentiation workload: d<E>/dJ2 of the
    triangular J1-J2 lattice (J1=1, J2=0.1, L=96, S=1) at T=0.1 from the
    covariance of E and the adjoint dE/dJ2 over 128 chains x 80 samples,
    beside a central finite difference of <E> from runs at J2 +- 0.005 with
    the same seeds.  Each run anneals to T (22 temperatures x 100 sweeps),
    then 200 + 80 x 10 cone-move Metropolis sweeps, each after 5
    overrelaxation sweeps."""
    from csmc.thermal import thermal_coupling_gradient

    def run(j2):
        uc = W.triangular_j1j2(1.0, j2)
        w = np.zeros(uc.n_params())
        for t in range(6, 12):
            w[9 * t + np.array([0, 4, 8])] = 1.0
        b = csmc.SpinBatch([csmc.Lattice((L, L), uc, 1.0, seed=5)] * n, precision=64)
        b.randomize(7)
        b.set_counters(0)
        b.anneal(W.anneal_temperatures(1.0, 0.9, T), 100, 5, csmc.CONE_ADAPTIVE, seed=11)
        out = thermal_coupling_gradient(b, T, samples, sweeps_between=10, overrelaxation=5, thermalize=200,
                                        alg=csmc.CONE, sigma=0.3, seed=11, directions=w)
        b.close()
        return out

    t0 = time.perf_counter()
    g, err, m = run(J2)
    secs = time.perf_counter() - t0
    mp, mm = run(J2 + d)[2], run(J2 - d)[2]
    fd = (mp["energy_per_system"] - mm["energy_per_system"]) / (2 * d)
    return {"value": round(float(g), 2), "err": round(float(err), 2),
            "finite_difference": round(float(fd.mean()), 2),
            "finite_difference_err": round(float(fd.std(ddof=1) / np.sqrt(n)), 2),
            "T": T, "J2": J2, "samples": m["samples"], "seconds": round(secs, 3),
            "T0_limit": 3 * L * L}


PEAK_FP64_TFLOPS = 78.6
SSF_FLOPS_PER_TERM = 12  # 3 components x (cos, sin) x one multiply-add


def ssf_cpu_baseline(pos, spins, ks, seconds, threads):
    """The oracle's compute_equal_time_correlations on the host: the momenta
    split over `threads` host threads (the C loop runs without the GIL), one
    system after another until about `seconds` have passed.  Returns
    terms/s (momenta x sites per system) and a description of the sample."""
    orc = baseline_oracle()
    threads = max(1, threads)
    parts = np.array_split(np.arange(ks.shape[1]), threads)
    stop = time.perf_counter() + seconds
    done = [0] * threads
    systems = [0] * threads

    def work(k):
        kp = np.ascontiguousarray(ks[:, parts[k]])
        r = 0
        while time.perf_counter() < stop:
            orc.equal_time_correlations(pos, spins[r % len(spins)], kp)
            done[k] += kp.shape[1] * pos.shape[1]
            systems[k] += 1
            r += 1

    t0 = time.perf_counter()
    pool = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    dt = time.perf_counter() - t0
    sample = (f"oracle compute_equal_time_correlations (FP64 sincos per term), {ks.shape[1]} momenta split "
              f"over {threads} host threads, {min(systems)}-{max(systems)} systems per thread in {dt:.1f} s")
    return sum(done) / dt, sample


def bench_ssf(args, csmc, W):
    """Equal-time structure factor S^ab(k) of 64 FP32 replicas of the
    honeycomb L = 64 lattice (8192 sites) at the 65 x 65 allowed momenta
    (get_allowed_wavevectors, one zone copy): one call per step, inputs and
    outputs resident in HBM for `value`; the output copy to the host is
    reported beside it."""
    from csmc.reciprocal import get_allowed_wavevectors
    L, R = args.L, args.replicas
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice((L, L), uc, 0.5, seed=3)
    ks = get_allowed_wavevectors(uc, (L, L))
    nk, N = ks.shape[1], lat.size
    ctx = csmc.default_context()
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(3)
    steps = min(args.steps, 50)
    warm = max(2, min(args.warmup, 5))
    for _ in range(warm):
        b.spin_correlations(ks)
    ctx.sync()
    ctx.timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.spin_correlations(ks)
    dt = time.perf_counter() - t0
    kern_ms, launches = ctx.timing_read()
    ctx.timing(False)
    terms = nk * N * R
    avg_ms = kern_ms / launches
    value = terms / (avg_ms * 1e-3)
    flops = SSF_FLOPS_PER_TERM * terms
    achieved = flops / (avg_ms * 1e-3) / 1e12
    kname = "rocBLAS dgemm path" if os.environ.get("CSMC_SSF") == "blas" else "k_ssf_mfma"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_ssf_mfma_honeycomb_L64_R64_traffic.json")
    if kname == "k_ssf_mfma" and L == 64 and R == 64 and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
    out = {"metric": "equal-time structure-factor terms/s (momenta x sites x systems) on 1 MI355X; "
                     "% of FP64 MFMA roofline",
           "value": value, "unit": "terms/s", "n_gpus": 1, "steps": steps, "warmup": warm,
           "ms_per_step": avg_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp64", "data": "synthetic",
           "config": {"workload": f"compute_equal_time_correlations of {R} replicas of the honeycomb L={L} "
                                  f"lattice ({N} sites, FP32 spins) at the {nk} allowed momenta of one zone "
                                  "copy, all replicas in one call",
                      "sites": N, "systems": R, "momenta": nk},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                        "kernel": kname, "flops_per_launch": flops, "avg_launch_ms": round(avg_ms, 4)},
           "with_host_copy": {"value": terms * steps / dt, "unit": "terms/s",
                              "note": "wall clock per call, including the output copy to the host "
                                      f"({R * nk * 9 * 8 / 1e6:.1f} MB over PCIe)"}}
    if not args.no_cpu:
        spins = [b.get_spins(r) for r in range(4)]
        cpu, sample = ssf_cpu_baseline(lat.site_positions, spins, ks, args.cpu_seconds, args.cpu_threads)
        out["cpu_baseline"] = {"value": cpu, "unit": "terms/s", "cores": args.cpu_threads, "kind": "port",
                               "sample": sample, "gpu_over_cpu": value / cpu}
    b.close()
    log(f"[bench] ssf: {avg_ms:.3f} ms per call, {achieved:.2f} TFLOP/s")
    finish(out, args)


if __name__ == "__main__":
    main()
