"""Train the learned kernel cost model on this device and compare the fusers.

    python -m bohrium_amd.costmodel_train [--out PATH] [--train 100:130] [--test 1000:1012]

1. Collect: run seeded random array programs (workloads.random_dag, one flush each) under the
   analytic greedy fuser, no fusion and the serial fuser with launch profiling on; every jitk
   kernel gives one sample (its block features, its device microseconds).
2. Fit: standardise the features and train the MLP (bohrium_amd.costmodel) on the GPU, 80/20
   split, and report the held-out error of log(microseconds).
3. Evaluate: install the model and run held-out programs under every fuser, including "learned",
   reporting the device time of each and the per-program best.

The model shipped in bohrium_amd/data/ was made this way on an MI355X.
"""
import argparse
import os
import time

import numpy as np

from . import _bh_api as api
from . import costmodel
from . import workloads

PIPELINES = {
    "greedy": "greedy, push_reductions_inwards, collapse_redundant_axes",
    "none": "push_reductions_inwards, collapse_redundant_axes",
    "serial": "serial, push_reductions_inwards, collapse_redundant_axes",
    "learned": "learned, push_reductions_inwards, collapse_redundant_axes",
}
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "costmodel_mi355x.npz")


def program_size(seed):
    """Instructions of random program `seed` (50..399)."""
    return 50 + (seed * 37) % 350


def _run_program(seed, sides):
    outs = workloads.random_dag(seed, program_size(seed), sides=sides)
    from .bharray import asnumpy_many
    asnumpy_many([o[:1] if o.ndim == 1 else o[:1, :1] for o in outs])


def run_profiled(seed, pipeline, reps=1, sides=(512, 1024, 2048)):
    """(device ms, launch text) of program `seed` under `pipeline`: the best of `reps` profiled runs
    after one warm-up run that compiles the kernels."""
    api.message("fuser_list: " + PIPELINES[pipeline])
    _run_program(seed, sides)
    best, text = None, ""
    for _ in range(reps):
        api.message("profile: on")
        api.message("launches")
        _run_program(seed, sides)
        t = api.message("launches")
        api.message("profile: off")
        dev = sum(float(l.split()[3]) for l in t.splitlines() if l.strip())
        if best is None or dev < best:
            best, text = dev, t
    return best, text


def collect(seeds, pipelines=("greedy", "none", "serial"), sides=(512, 1024, 2048), log=print):
    X, Y = [], []
    for s in seeds:
        for p in pipelines:
            _, text = run_profiled(s, p, 1, sides)
            x, y = costmodel.profiled_kernels(text)
            X.append(x)
            Y.append(y)
        log("collected program %d (%d instructions)" % (s, program_size(s)))
    return np.concatenate(X), np.concatenate(Y)


def evaluate(seeds, pipelines=tuple(PIPELINES), sides=(512, 1024, 2048), reps=2, log=print):
    """Per program: {pipeline: device ms}; plus the totals."""
    rows = []
    for s in seeds:
        r = {p: run_profiled(s, p, reps, sides)[0] for p in pipelines}
        rows.append((s, r))
        log("program %d (%d instructions): %s" % (s, program_size(s), {k: round(v, 3) for k, v in r.items()}))
    tot = {p: sum(r[p] for _, r in rows) for p in pipelines}
    best = sum(min(r.values()) for _, r in rows)
    return rows, tot, best


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--train", default="100:130")
    ap.add_argument("--test", default="1000:1012")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--eval-only", action="store_true", help="evaluate the model at --out without training")
    a = ap.parse_args(argv)
    api.set_flush_threshold(10 ** 7)
    if a.eval_only:
        costmodel.CostModel.load(a.out).install()
        _report(a)
        return
    lo, hi = (int(v) for v in a.train.split(":"))
    t0 = time.time()
    X, Y = collect(range(lo, hi))
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
    model.install()
    _report(a)


def _report(a):
    lo, hi = (int(v) for v in a.test.split(":"))
    _, tot, best = evaluate(range(lo, hi))
    print("total device ms", {k: round(v, 3) for k, v in tot.items()}, "per-program best %.3f" % best)
    print("learned/greedy %.3f learned/best %.3f greedy/best %.3f serial/best %.3f"
          % (tot["learned"] / tot["greedy"], tot["learned"] / best, tot["greedy"] / best, tot["serial"] / best))
    api.set_flush_threshold(1000)
    api.message("fuser_list: " + PIPELINES["greedy"])


if __name__ == "__main__":
    main()
