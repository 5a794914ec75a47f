"""Benchmark: effective HBM GB/s of the fused Black-Scholes workload on the hip engine.

Workload (one step = one full run of the configuration): S ~ U[58,62], K ~ U[55,65],
T ~ U[0.25,2] float64, N = 2e7 (generated on the device by philox before timing);
10 time steps, each a fused map (A&S polynomial CND with exp/log/sqrt) + sum.  The
time steps are one do_while body replayed by the engine.  Steps are issued back to
back (the engine's stream is asynchronous); the timed region ends with one
synchronizing fetch of every step's 10 sums.
Algorithmic bytes per time step = read S, K, T once = 24 * N bytes.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--no-secondary]

Prints one JSON line with the value (GB/s over the whole timed job), the roofline
of the dominant kernel (per-launch device time from HIP events on the engine
stream), the CPU baseline (the oracle's fused OpenMP form on the host cores) and
secondary measurements of the 2-D Jacobi and batched-arrays workloads and of the MFMA GEMM.
"""
import argparse
import faulthandler
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TIME_STEPS = 10


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    # 30 timed runs (about 50 ms): the host queues a run in a third of its device time, so after the
    # first run's start-up (about 0.7 ms) the device stays busy; fewer runs weigh the start-up in more.
    p.add_argument("--steps", type=int, default=30)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=20_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-n", type=int, default=20_000_000)
    p.add_argument("--no-secondary", action="store_true")
    p.add_argument("--cpu-only", action="store_true",
                   help="print only the CPU baseline as JSON (the bench runs it this way in a child process)")
    p.add_argument("--gemm", action="store_true",
                   help="also time blas_gemm on 8192^2 operands (python -m bohrium_amd.blas does the same)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a plain 256 MiB device copy 5 times (counter calibration for rocprofv3 --pmc)")
    return p.parse_args()


def traffic_from_profiles():
    """HBM bytes per launch of the Black-Scholes kernel measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), if recorded."""
    path = os.path.join(ROOT, "profiles", "black_scholes_step_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def host_cpu():
    """CPU model, logical CPUs and physical cores of this host (from /proc/cpuinfo)."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k:
                    if core is not None:
                        cores.add((phys, core))
                    phys = core = None
        if core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(cores) or None,
            "usable_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_baseline(n):
    import bohrium_amd as bh
    from bohrium_amd import workloads, _bh_api as api
    from oracle import omp_ve
    with api.record_only() as r:
        S, K, T = workloads.black_scholes_inputs(n)
        bh.flush()
        sums = [workloads.black_scholes_step(S, K, T, i) for i in range(TIME_STEPS)]
        bh.flush()
        del sums, S, K, T
    prog = omp_ve.FusedProgram(r.trace)
    prog.run_next_group()           # inputs
    t = prog.run_next_group()       # the 10 time steps
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(24.0 * n * TIME_STEPS / t / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port", "seconds_per_run": round(t, 4), "host": host_cpu(),
            "sample": "full workload at N=%d (%d time steps), oracle fused OpenMP C99 (gcc -O3 -march=native "
                      "-fopenmp), %d threads" % (n, TIME_STEPS, threads)}


def cpu_secondary_baselines():
    """The oracle's fused OpenMP form on bounded samples of the other workloads (same host, same
    threads): Monte-Carlo pi at N = 1e7, 10 Jacobi sweeps of the 8192^2 grid, and the first 512 of
    the 4096 batched arrays.  Reported per unit of the GPU measurement."""
    import numpy as np
    import bohrium_amd as bh
    from bohrium_amd import workloads, _bh_api as api
    from oracle import omp_ve
    out = {}
    with api.record_only() as r:
        c = workloads.mc_pi(10_000_000)
        bh.flush()
        del c
    prog = omp_ve.FusedProgram(r.trace)
    t = prog.run_next_group()
    out["mc_pi"] = {"value": round(1e7 / t, 1), "unit": "points/s", "seconds": round(t, 4),
                    "sample": "full workload: N=1e7 points"}
    sweeps = 10
    with api.record_only() as r:
        G = workloads.jacobi_grid(8192)
        bh.flush()
        for _ in range(sweeps):
            d = workloads.jacobi_step(G)
            bh.flush()
            del d
        del G
    prog = omp_ve.FusedProgram(r.trace)
    prog.run_next_group()
    t = sum(prog.run_next_group() for _ in range(sweeps))
    alg = 8192 * 8192 * 4 + 8190 * 8190 * 4
    out["jacobi"] = {"value": round(alg * sweeps / t / 1e9, 3), "unit": "GB/s", "ms_per_sweep": round(t / sweeps * 1e3, 3),
                     "sample": "%d sweeps of the 8192x8192 float32 grid (stencil + max|delta| + copy-back, "
                               "the program as written)" % sweeps}
    sizes = workloads.batched_sizes(4096)[:512]
    with api.record_only() as r:
        xs, idxs = workloads.batched_inputs(sizes)
        bh.flush()
        res = workloads.batched_step(xs, idxs)
        bh.flush()
        del res, xs, idxs
    prog = omp_ve.FusedProgram(r.trace)
    prog.run_next_group()
    t = prog.run_next_group()
    elems = int(np.sum(sizes))
    out["batched"] = {"value": round(32.0 * elems / t / 1e9, 3), "unit": "GB/s", "seconds": round(t, 4),
                      "sample": "the first 512 of the 4096 arrays (%d elements), 32 B/element" % elems}
    return out


def cpu_baseline_child(n, others=True):
    """The CPU baseline in a child process of its own: its OpenMP threads do not share the process
    with the GPU runtime, and a failure there leaves the bench line intact (cpu_baseline then
    carries the error)."""
    import subprocess
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-only", "--cpu-n", str(n)] + ([] if others else ["--no-secondary"])
    r = subprocess.run(cmd, env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=900)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"value": None, "unit": "GB/s", "cores": None, "kind": "port",
                "error": "child exited with %d: %s" % (r.returncode, r.stderr[-300:])}
    return json.loads(lines[-1])


def autotune_trials(bh):
    for line in bh.message("statistic").splitlines():
        if line.startswith("autotune_trials:"):
            return float(line.split(":")[1])
    return 0.0


def settle(bh, run, limit):
    """Calls run() until a call starts no autotuner trial launch (at most `limit` calls), then settles
    the finished trials; returns the number of calls."""
    for k in range(1, limit + 1):
        before = autotune_trials(bh)
        run()
        if autotune_trials(bh) == before:
            break
    bh.message("autotune")
    return k


def launches(bh):
    return [l.split() for l in bh.message("launches").splitlines() if l.strip()]


def mc_pi_secondary(bh, workloads):
    """Monte-Carlo pi at N = 1e7: philox uniforms fused with the test and the int64 count (one
    FULL_REDUCE launch and its finalize); compute-bound, reported as points per second."""
    n, reps = 10_000_000, 20
    count = int(workloads.mc_pi(n).asnumpy())
    settle(bh, lambda: workloads.mc_pi(n).asnumpy(), 100)  # the autotuner's trials of the launch shape
    bh.message("profile: on")
    launches(bh)
    for _ in range(3):
        workloads.mc_pi(n).asnumpy()
    ls = launches(bh)
    bh.message("profile: off")
    red = [float(l[3]) for l in ls if l[0] == "reduce" and int(l[1]) == n]
    kernel_ms = sum(red) / max(1, len(red))
    t0 = time.perf_counter()
    for _ in range(reps):
        workloads.mc_pi(n).asnumpy()
    wall = (time.perf_counter() - t0) / reps
    return {"workload": "Monte-Carlo pi, N=1e7 points from philox2x32-10 (key 42), int64 count",
            "count": count, "pi_estimate": 4.0 * count / n, "kernel_ms": round(kernel_ms, 4),
            "points_per_s_kernel": round(n / (kernel_ms * 1e-3), 1) if kernel_ms > 0 else None,
            "ms_per_evaluation_wall": round(wall * 1e3, 4)}


def jacobi_secondary(bh, workloads):
    """2-D Jacobi fp32 8192^2, 200 sweeps replayed as one graph per sweep."""
    n, iters = 8192, 200
    G = workloads.jacobi_grid(n)
    # Warm-up in 40-sweep runs until a run starts no autotuner trial (at most 8 runs), then one more
    # run for the repeat's graphs of the chosen launch shapes.
    settle(bh, lambda: workloads.jacobi_repeat(G, 40).asnumpy(), 8)
    workloads.jacobi_repeat(G, 40).asnumpy()
    bh.message("profile: on")
    launches(bh)
    workloads.jacobi_repeat(G, 3).asnumpy()
    ls = launches(bh)
    bh.message("profile: off")
    red = [float(l[3]) for l in ls if l[0] == "reduce"]
    mp = [float(l[3]) for l in ls if l[0] == "map" and int(l[1]) > 1]
    before = int(bh.message("graph replays"))
    t0 = time.perf_counter()
    d = workloads.jacobi_repeat(G, iters).asnumpy()
    dt = time.perf_counter() - t0
    replays = int(bh.message("graph replays")) - before
    alg = n * n * 4 + (n - 2) * (n - 2) * 4
    return {"workload": "jacobi fp32 %dx%d, %d sweeps (c[...] = stencil(c) with max|delta|: the stencil writes the "
                        "grid's twin buffer and the buffers are exchanged), do_while replayed as a HIP graph"
                        % (n, n, iters),
            "ms_per_sweep": round(dt * 1e3 / iters, 4), "effective_GBs": round(alg * iters / dt / 1e9, 1),
            "algorithmic_bytes_per_sweep": alg, "graph_replays": replays,
            "kernel_ms": {"stencil+max (reduce)": round(sum(red) / max(1, len(red)), 4),
                          "other maps": round(sum(mp) / max(1, len(mp)), 4)},
            "last_delta": float(d)}


def batched_secondary(bh, workloads, api):
    """4096 uneven float32 arrays (256..262144 elements) in one flush: batched launches level by level
    (the default), the whole schedule as one persistent launch, and one launch per block."""
    sizes = workloads.batched_sizes(4096)
    elems = int(sizes.sum())
    api.set_flush_threshold(10 ** 7)
    xs, idxs = workloads.batched_inputs(sizes)
    bh.flush()
    out = {"workload": "4096 float32 arrays, sizes floor(exp(U[ln 256, ln 262144])) (seed 42, %d elements): "
                       "tanh(0.5x+0.1), sum, int64 cumsum(x>0), y[idx]" % elems,
           "algorithmic_bytes": 32 * elems}

    def run():
        sums, cums, gathered = workloads.batched_step(xs, idxs)
        bh.asnumpy_many([sums[-1], cums[-1][:1], gathered[-1][:1]])

    for label, bm, persistent in (("batched", 8, "off"), ("persistent", 8, "on"), ("one_launch_per_block", 0, "off")):
        bh.message("batch_min: %d" % bm)
        bh.message("batch_persistent: " + persistent)
        run()
        t0 = time.perf_counter()
        run()
        dt = time.perf_counter() - t0
        bh.message("profile: on")
        launches(bh)
        run()
        ls = launches(bh)
        bh.message("profile: off")
        dev = sum(float(l[3]) for l in ls)
        out[label] = {"launches": len(ls), "device_ms": round(dev, 3), "wall_ms": round(dt * 1e3, 1),
                      "device_GBs": round(32.0 * elems / (dev * 1e-3) / 1e9, 1) if dev > 0 else None}
    bh.message("batch_min: 8")
    bh.message("batch_persistent: off")
    api.set_flush_threshold(1000)
    return out


def streaming_secondary(bh):
    """Memory-bound fused blocks on 2^28 float32 elements (1 GiB): a map + sum, a map, and a max of |x - c|,
    each after the autotuner's trials have settled (device time of the launches from HIP events)."""
    import numpy as np
    n = 1 << 28
    x = bh.random.RandomState(1).random_sample(n).astype(np.float32)
    x[:1].asnumpy()
    out = {"workload": "float32 x of 2^28 elements (1 GiB), philox uniforms"}
    cases = (("sum(x*2+1)", lambda: (x * 2.0 + 1.0).sum(), 4 * n),
             ("z = x*2+1", lambda: x * 2.0 + 1.0, 8 * n),
             ("max|x-0.5|", lambda: bh.ufunc("maximum").reduce(bh.ufunc("absolute")(x - 0.5), 0), 4 * n))
    for label, fn, nbytes in cases:
        def run():
            r = fn()
            (r[:1] if r.ndim else r).asnumpy()
        settle(bh, run, 100)
        run()
        bh.message("profile: on")
        launches(bh)
        ts = []
        for _ in range(5):
            run()
            ts.append(sum(float(l[3]) for l in launches(bh) if l[0] in ("reduce", "map")))
        bh.message("profile: off")
        t = min(ts)
        out[label] = {"device_ms": round(t, 4), "algorithmic_bytes": nbytes, "GBs": round(nbytes / t / 1e6, 1),
                      "frac_of_hbm_peak": round(nbytes / t / 1e6 / HBM_PEAK_GBS, 4)}
    del x
    return out


def gemm_secondary(bh):
    """blas_gemm (the MFMA extension method) on 8192 x 8192 operands, fp64 and fp32: wall time per
    product between host syncs, operands resident."""
    import numpy as np
    out = {}
    n = 8192
    for name, dt, peak in (("f64", np.float64, 78.6), ("f32", np.float32, 157.3)):
        R = bh.random.RandomState(7)
        a = R.random_sample((n, n)).astype(dt)
        b = R.random_sample((n, n)).astype(dt)
        bh.asnumpy_many([a[:1, :1], b[:1, :1]])
        def product():
            # one element of C copied back (a view's asnumpy would copy all of C's base)
            return (bh.blas.gemm(a, b)[:1, :1] * 1.0).asnumpy()
        product()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            product()
            ts.append(time.perf_counter() - t0)
        tf = 2.0 * n ** 3 / min(ts) / 1e12
        out[name] = {"n": n, "ms": round(min(ts) * 1e3, 3), "TFLOPs": round(tf, 1), "frac_of_mfma_peak": round(tf / peak, 3)}
        del a, b
    return out


def main():
    faulthandler.enable()  # a fatal signal prints the Python stack of every thread
    args = parse()
    if args.cpu_only:
        base = cpu_baseline(args.cpu_n)
        if not args.no_secondary:
            base["others"] = cpu_secondary_baselines()
        print(json.dumps(base))
        return
    if args.gpus != 1:
        print("bench.py: the engine runs on one GPU; --gpus %d is not supported (a sharded form would give each "
              "rank a contiguous slice of every array; it is designed, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import bohrium_amd as bh
    from bohrium_amd import workloads, _bh_api as api

    # Launch shapes are tuned by this run's own warm-up: choices another process left in the kernel
    # cache's table (a test run's noisy trials) are not taken over.
    bh.message("autotune: forget")
    n = args.n
    S, K, T = workloads.black_scholes_inputs(n)
    bh.asnumpy_many([S[:1], K[:1], T[:1]])  # inputs resident in HBM before timing

    def run():
        # The 10 time steps are one recorded loop body replayed by the engine (do_while with a
        # sliding view for the time offset and the output slot); nothing is fetched here.
        return workloads.black_scholes_repeat(S, K, T, TIME_STEPS)

    for _ in range(args.warmup):
        run().asnumpy()
    # Warm-up goes on (at most 20 more runs) until a run starts no autotuner trial, then once more
    # so the replay graphs of the chosen launch shapes exist: the timed runs are the steady state.
    settle_runs = 1
    for _ in range(20):
        before = autotune_trials(bh)
        run().asnumpy()
        settle_runs += 1
        bh.message("autotune")
        if autotune_trials(bh) == before:
            break
    run().asnumpy()
    t0 = time.perf_counter()
    outs = [run() for _ in range(args.steps)]
    vals = bh.asnumpy_many(outs)
    elapsed = time.perf_counter() - t0
    bytes_per_run = 24.0 * n * TIME_STEPS
    value = bytes_per_run * args.steps / elapsed / 1e9
    del outs

    # Per-launch device time of the dominant kernel (HIP events on the engine stream): the
    # main kernel and its one-workgroup finalize kernel together.
    bh.message("profile: on")
    launches(bh)
    for _ in range(5):
        run().asnumpy()
    ls = launches(bh)
    bh.message("profile: off")
    main_k = [float(l[3]) for l in ls if l[0] == "reduce" and int(l[1]) == n]
    avg_ms = sum(main_k) / len(main_k) if main_k else float("nan")
    achieved = 24.0 * n / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else float("nan")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                "kernel": "black-scholes step: fused map+sum (FULL_REDUCE main + finalize)",
                "avg_kernel_ms": round(avg_ms, 4), "launches_measured": len(main_k),
                "algorithmic_bytes_per_launch": int(24 * n)}

    if args.calibrate_copy:
        a = bh.random.RandomState(1).random_sample(32 * 1024 * 1024)
        b = bh.empty(32 * 1024 * 1024)
        for _ in range(5):
            b[...] = a
            bh.asnumpy_many([b[:1]])
        del a, b

    secondary = None
    if not args.no_secondary:
        del S, K, T
        secondary = {"mc_pi": mc_pi_secondary(bh, workloads), "jacobi": jacobi_secondary(bh, workloads),
                     "batched": batched_secondary(bh, workloads, api), "streaming": streaming_secondary(bh)}
        if args.gemm:
            secondary["blas_gemm"] = gemm_secondary(bh)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_child(args.cpu_n, others=not args.no_secondary)
        others = cpu.pop("others", None) if cpu else None
        if secondary is not None and others:
            for k, v in others.items():
                if k in secondary:
                    secondary[k]["cpu_baseline"] = v

    out = {"metric": "effective HBM GB/s (% of peak), s/run vs OpenMP ve; learned vs greedy fuser",
           "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "warmup_autotune_settle_runs": settle_runs,
           "ms_per_step": round(elapsed / args.steps * 1000, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (philox, seed 42)",
           "config": {"workload": "black_scholes fp64 N=%d, %d time steps, call-price sum per step" % (n, TIME_STEPS),
                      "n": n, "time_steps": TIME_STEPS, "bytes_per_step": int(bytes_per_run),
                      "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4)},
           "roofline": roofline, "cpu_baseline": cpu,
           "seconds_per_run": round(elapsed / args.steps, 5),
           "speedup_vs_cpu_baseline": (round(cpu["seconds_per_run"] / (elapsed / args.steps), 1) if cpu and cpu.get("seconds_per_run") else None),
           "secondary": secondary,
           "autotune": {k: float(v) for k, v in (l.split(":") for l in bh.message("statistic").splitlines()
                                                  if l.startswith("autotune_"))},
           "check": {"first_step_sum": float(vals[0][0]), "last_step_sum": float(vals[-1][-1])}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
