"""Train the launch-shape model of the autotuner on this device and judge it against measurement.

    python -m bohrium_amd.launchmodel_train [--out PATH] [--train 100:112] [--test 1000:1006]

1. Collect: run programs whose kernels have at least 2^20 elements (seeded random array programs
   at 1024^2 and 2048^2, Black-Scholes, Monte-Carlo pi, Jacobi sweeps, scans, row sums) with the
   autotuner measuring every candidate launch shape (2 timed launches each) until their kernels
   settle; every timed launch is one sample: the kernel's block features, the candidate's launch
   features, its device microseconds ("autotune samples").
2. Fit the MLP (bohrium_amd.costmodel, 19 inputs) on the GPU, 80/20 split; report the held-out error
   of log(microseconds).
3. Judge: install it as the launch-shape model and tune held-out programs by measurement again.
   Each settled trial reports every candidate's measured time and the model's prediction
   ("autotune outcomes"), so the model's choice is compared with the measured-best candidate and
   with the configured knobs on the same kernels.

The model shipped in bohrium_amd/data/ was made this way on an MI355X.
"""
import argparse
import os
import time

import numpy as np

from . import _bh_api as api
from . import costmodel
from . import workloads

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "launchmodel_mi355x.npz")


def _stat(name):
    for line in api.message("statistic").splitlines():
        if line.startswith(name + ":"):
            return float(line.split(":")[1])
    return 0.0


def _settle(run, max_runs=40):
    """Run `run()` until a run launches no trial (its kernels have settled)."""
    for it in range(max_runs):
        before = _stat("autotune_trials")
        run()
        api.message("autotune")  # reads finished measurements, settles trials
        if it > 0 and _stat("autotune_trials") == before:
            return it + 1
    return max_runs


def _program_runs(seed, side):
    from .bharray import asnumpy_many

    def run():
        outs = workloads.random_dag(seed, 60 + (seed * 37) % 120, sides=(side,))
        asnumpy_many([o[:1] if o.ndim == 1 else o[:1, :1] for o in outs])
    return run


def _fixed_runs():
    """Kernels of the benchmark workloads at several sizes."""
    import bohrium_amd as bh
    runs = []
    for n in (1 << 20, 1 << 22, 1 << 24):
        S, K, T = workloads.black_scholes_inputs(n)
        runs.append(lambda S=S, K=K, T=T: workloads.black_scholes_step(S, K, T, 0).asnumpy())
        runs.append(lambda n=n: workloads.mc_pi(n).asnumpy())
        x = bh.random.RandomState(3).random_sample(n)
        runs.append(lambda x=x: (x > 0.5).astype(np.int64).cumsum()[-1:].asnumpy())
    for side in (1024, 2048, 4096):
        G = workloads.jacobi_grid(side)
        runs.append(lambda G=G: workloads.jacobi_step(G).asnumpy())
        a = bh.random.RandomState(4).random_sample((side, side)).astype(np.float32)
        runs.append(lambda a=a: (a * 2.0 + 1.0).sum(axis=1)[:1].asnumpy())
    return runs


def collect(seeds, sides=(1024, 2048), log=print):
    api.message("autotune: on")
    api.message("autotune_mode: measure")
    api.message("autotune_samples: 2")
    api.message("autotune: forget")
    api.message("autotune samples")  # drop earlier samples
    runs = [_program_runs(s, side) for s in seeds for side in sides] + _fixed_runs()
    X, Y = [], []
    for k, run in enumerate(runs):
        n = _settle(run)
        x, y = costmodel.tuning_samples(api.message("autotune samples"))
        X.append(x)
        Y.append(y)
        log("program %d/%d: %d runs, %d samples" % (k + 1, len(runs), n, len(y)))
    api.message("autotune_samples: 3")
    return np.concatenate(X), np.concatenate(Y)


def judge(seeds, sides=(1024, 2048), extra=True, log=print):
    """Tune held-out programs by measurement with the installed launch model; returns the settled
    trials' outcomes and the totals (measured ms) of the model's picks, the measured-best
    candidates and the configured knobs."""
    api.message("autotune: on")
    api.message("autotune_mode: measure")
    api.message("autotune_samples: 2")
    api.message("autotune: forget")
    api.message("autotune outcomes")
    runs = [_program_runs(s, side) for s in seeds for side in sides] + (_fixed_runs() if extra else [])
    for run in runs:
        _settle(run)
    out = costmodel.tuning_outcomes(api.message("autotune outcomes"))
    api.message("autotune_samples: 3")
    picks = [o for o in out if o[3] is not None]
    tot_model = sum(ms[int(np.argmin(pred))] for _, _, ms, pred in picks)
    tot_best = sum(min(ms) for _, _, ms, _ in picks)
    tot_default = sum(ms[0] for _, _, ms, _ in picks)
    log("%d kernels: model picks %.3f ms, measured best %.3f ms, configured knobs %.3f ms"
        % (len(picks), tot_model, tot_best, tot_default))
    for k in (2, 3):
        tot_k = sum(min(ms[0], min(ms[j] for j in top_k(pred, k))) for _, _, ms, pred in picks)
        trials = sum(min(len(ms), k + 1) for _, _, ms, _ in picks)
        log("  prune to the configured knobs + the model's best %d: %.3f ms (%d candidates measured, of %d)"
            % (k, tot_k, trials, sum(len(ms) for _, _, ms, _ in picks)))
    return picks, tot_model, tot_best, tot_default


def top_k(pred, k):
    """Indices of the k candidates with the lowest predicted time."""
    return [int(j) for j in np.argsort(np.asarray(pred))[:k]]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--train", default="100:112")
    ap.add_argument("--test", default="1000:1006")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--eval-only", action="store_true")
    a = ap.parse_args(argv)
    api.set_flush_threshold(10 ** 7)
    lo, hi = (int(v) for v in a.test.split(":"))
    if a.eval_only:
        model = costmodel.CostModel.load(a.out)  # (installed only while this object lives)
        model.install_launch()
        judge(range(lo, hi), extra=False)
        return
    tlo, thi = (int(v) for v in a.train.split(":"))
    t0 = time.time()
    X, Y = collect(range(tlo, thi))
    print("samples %d collected in %.1f s" % (len(Y), time.time() - t0), flush=True)
    perm = np.random.RandomState(0).permutation(len(Y))
    ntr = int(0.8 * len(Y))
    tr, va = perm[:ntr], perm[ntr:]
    model, losses = costmodel.fit(X[tr], Y[tr], hidden=32, steps=a.steps, lr=3e-3)
    err = model.predict(X[va]) - np.log(np.maximum(Y[va], 1e-3))
    base = np.log(np.maximum(Y[va], 1e-3)) - np.log(np.maximum(Y[tr], 1e-3)).mean()
    print("loss %.4f -> %.4f; held-out rmse(log us) %.3f (predicting the mean: %.3f)"
          % (losses[0], losses[-1], np.sqrt(np.mean(err ** 2)), np.sqrt(np.mean(base ** 2))), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    model.save(a.out)
    model.install_launch()
    judge(range(lo, hi), extra=False)
    api.set_flush_threshold(1000)


if __name__ == "__main__":
    main()
