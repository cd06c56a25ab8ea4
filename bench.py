"""Benchmark of the sweep hot path on one MI355X.

Workload: honeycomb Kitaev-Heisenberg model (J = 1, K = -2 on the x/y/z
bonds), L = 64 (8192 spins), S = 1/2, parallel tempering with 64 replicas at
geometric temperatures 0.01 .. 1.0, exchange every 10 sweeps, one
Metropolis sweep every 10 overrelaxation sweeps; FP32 spin updates, FP64
energies, exchange decisions and observable binning.

One step = one exchange period: 10 overrelaxation sweeps + 1 Metropolis
sweep + FP64 total energy of every replica + the replica exchange, i.e.
11 * 64 * 8192 = 5,767,168 spin updates.  Metric: spin-updates/s.

Beside the batched FP32 form it measures the plain FP64 GPU form (one
replica at a time, one launch per colour pass) on the same lattice and the
oracle (CPU FP64 restatement of the reference's PT loop: random-site
Metropolis, sequential overrelaxation) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "classicalspinmc.jl_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
BYTES_PER_UPDATE_FP32 = 24  # read + write own spin, 3 x 4 B each


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run device-to-device copies of this many bytes (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU form")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; --gpus N > 1 is not "
                                   "supported (replica batches are sharded by process, see DESIGN.md)"}))
        sys.exit(2)

    import csmc
    from csmc import workloads as W

    L, R = args.L, args.replicas
    uc = W.honeycomb_kitaev(1.0, -2.0)
    lat = csmc.Lattice((L, L), uc, 0.5, seed=2)
    N = lat.size
    temps = W.geometric_temperatures(0.01, 1.0, R)
    sweeps_per_step = 10
    params = dict(t_thermalization=10 ** 4, t_measurement=10 ** 5, probe_rate=10, swap_rate=10,
                  overrelaxation_rate=10, report_interval=10 ** 9, checkpoint_rate=10 ** 9)
    ctx = csmc.default_context()
    b = csmc.SpinBatch([lat] * R, precision=32, ctx=ctx)
    b.randomize(2)
    b.pt_init(temps, params, csmc.METROPOLIS, seed=2)
    log(f"[bench] lattice N={N} replicas={R} colours={lat.info.n_colours}")

    # warmup (thermalization sweeps)
    b.pt_advance(args.warmup * sweeps_per_step)
    ctx.sync()
    log(f"[bench] warmup {args.warmup} steps done")

    # timed region: K steps = K exchange periods
    ctx.timing(True)
    start_sweep = args.warmup * sweeps_per_step
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
        f"{launches} launches, {kern_ms / max(launches, 1):.3f} ms avg")

    # sanity: FP64 energies finite, exchanges happened, spins normalised
    st = b.pt_state()
    assert np.all(np.isfinite(st["E"])) and st["accepted_exchange"].sum() > 0
    s0 = b.get_spins(0)
    assert np.allclose(np.linalg.norm(s0, axis=1), 0.5, atol=1e-4)

    # roofline of the dominant kernel (the PT segment kernel, one launch per step)
    avg_ms = kern_ms / launches
    bytes_per_launch = BYTES_PER_UPDATE_FP32 * updates_per_step
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by
    # a copy kernel of known size) for this exact workload, see profiles/
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_wg_pt_honeycomb_L64_R64_traffic.json")
    if L == 64 and R == 64 and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "kernel": "k_wg (PT segment, workgroup per replica)",
                "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_ms, 5),
                "note": "spins and neighbour tables are LDS-resident during a launch: measured HBM "
                        "traffic is the per-launch load/store of the replicas, far below the "
                        "algorithmic 24 B per update"}

    out = {
        "metric": "spin-updates/s on 1 MI355X (batched FP32 PT vs plain); % of HBM roofline",
        "value": value, "unit": "spin-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"honeycomb Kitaev-Heisenberg J=1 K=-2, L={L} ({N} spins), S=1/2, "
                               f"parallel tempering, {R} replicas, T=0.01..1 geometric, "
                               "swap_rate=10, overrelaxation_rate=10",
                   "replicas": R, "sites": N, "updates_per_step": updates_per_step,
                   "precision": "fp32 spins, fp64 energies/exchange"},
        "roofline": roofline,
    }

    if not args.no_plain:
        # plain unbatched FP64 form: one replica at a time, one launch per
        # colour pass; one exchange period (10 OR + 1 Metropolis) on 4 replicas
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
        out["plain_gpu"] = {"value": plain, "unit": "spin-updates/s", "dtype": "fp64",
                            "form": "one replica at a time, one launch per colour pass",
                            "batched_over_plain": value / plain}
        pb.close()
        log(f"[bench] plain FP64 form: {plain / 1e9:.3f} G updates/s")

    if not args.no_cpu:
        import oracle.oracle as orc
        T = orc.Tables(uc, (L, L))
        rng = np.random.default_rng(0)
        v = rng.normal(size=(R, N, 3))
        spins = np.ascontiguousarray(0.5 * v / np.linalg.norm(v, axis=2, keepdims=True))
        E = np.zeros(R)
        perm = np.arange(R, dtype=np.intc)
        done, t0, sw = 0, time.perf_counter(), 0
        while time.perf_counter() - t0 < args.cpu_seconds and sw < 2000:
            done += orc.pt_reference_run(T, spins, temps, 0.5, sw, 10, 10, 10, 2, E, perm)
            sw += 10
        dtc = time.perf_counter() - t0
        cpu = done / dtc
        out["cpu_baseline"] = {"value": cpu, "unit": "spin-updates/s", "cores": 1, "kind": "port",
                               "sample": f"oracle PT loop (random-site Metropolis + sequential "
                                         f"overrelaxation + total_energy + exchange), {R} replicas x "
                                         f"{sw} sweeps of the same lattice, single thread",
                               "gpu_over_cpu": value / cpu}
        log(f"[bench] oracle CPU: {cpu / 1e6:.1f} M updates/s over {sw} sweeps")

    print(json.dumps(out))


if __name__ == "__main__":
    main()
