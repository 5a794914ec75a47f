"""Benchmark: effective HBM GB/s of the fused Black-Scholes workload on the hip engine.

Workload (one step = one full run of the configuration): S ~ U[58,62], K ~ U[55,65],
T ~ U[0.25,2] float64, N = 2e7 (generated on the device by philox before timing);
10 time steps, each a fused map (A&S polynomial CND with exp/log/sqrt) + sum.  The
time steps are one do_while body replayed by the engine; all 10 sums are fetched
once at the end of each run.
Algorithmic bytes per time step = read S, K, T once = 24 * N bytes.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]

Prints one JSON line with the value (GB/s over the whole timed job), the roofline
of the dominant kernel (per-launch device time from HIP events on the engine
stream) and the CPU baseline (the oracle's fused OpenMP form on the host cores).
"""
import argparse
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
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=20_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-n", type=int, default=20_000_000)
    return p.parse_args()


def traffic_from_profiles():
    """HBM bytes per launch of the Black-Scholes kernel measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), if recorded."""
    path = os.path.join(ROOT, "profiles", "black_scholes_step_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def cpu_baseline(n):
    import numpy as np
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
            "kind": "port", "seconds_per_run": round(t, 4),
            "sample": "full workload at N=%d (%d time steps), oracle fused OpenMP C99 (gcc -O3 -march=native "
                      "-fopenmp), %d threads" % (n, TIME_STEPS, threads)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (see DESIGN.md for the "
              "sharding design)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import bohrium_amd as bh
    from bohrium_amd import workloads

    n = args.n
    S, K, T = workloads.black_scholes_inputs(n)
    bh.asnumpy_many([S[:1], K[:1], T[:1]])  # inputs resident in HBM before timing

    def run():
        # The 10 time steps are one recorded loop body replayed by the engine (do_while with a
        # sliding view for the time offset and the output slot); one sync at the end.
        return workloads.black_scholes_repeat(S, K, T, TIME_STEPS).asnumpy()

    for _ in range(args.warmup):
        run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vals = run()
    elapsed = time.perf_counter() - t0
    bytes_per_run = 24.0 * n * TIME_STEPS
    value = bytes_per_run * args.steps / elapsed / 1e9

    # Per-launch device time of the dominant kernel (HIP events on the engine stream).
    bh.message("profile: on")
    bh.message("launches")
    for _ in range(max(2, args.steps)):
        run()
    lines = [l.split() for l in bh.message("launches").splitlines() if l.strip()]
    bh.message("profile: off")
    main_k = [float(l[3]) for l in lines if l[0] == "reduce" and int(l[1]) == n]
    avg_ms = sum(main_k) / len(main_k) if main_k else float("nan")
    achieved = 24.0 * n / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                "kernel": "black-scholes fused map+reduce (skeleton: reduce)", "avg_kernel_ms": round(avg_ms, 4),
                "launches_measured": len(main_k)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n)

    out = {"metric": "effective HBM GB/s (% of peak), s/run vs OpenMP ve; learned vs greedy fuser",
           "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed / args.steps * 1000, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (philox, seed 42)",
           "config": {"workload": "black_scholes fp64 N=%d, %d time steps, call-price sum per step" % (n, TIME_STEPS),
                      "n": n, "time_steps": TIME_STEPS, "bytes_per_step": int(bytes_per_run),
                      "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4)},
           "roofline": roofline, "cpu_baseline": cpu,
           "seconds_per_run": round(elapsed / args.steps, 5),
           "speedup_vs_cpu_baseline": (round(cpu["seconds_per_run"] / (elapsed / args.steps), 1) if cpu else None),
           "check": {"first_step_sum": float(vals[0]), "last_step_sum": float(vals[-1])}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
