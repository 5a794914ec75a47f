"""Write tests/golden/bhir/<name>.bhir: recorded flushes (the BhIR wire format of
bhc_record_flushes) of small versions of the benchmark programs, plus
<name>.json naming the bases whose contents the tests compare after a replay.

The programs run on a compile-only engine (BH_HIP_DRY_RUN=1, no GPU needed); the
recording holds the bytecode and the data of arrays created from host data.

    BH_HIP_DRY_RUN=1 python tests/golden/make_bhir_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "bhir")
sys.path.insert(0, ROOT)


def _mc_pi():
    from bohrium_amd import workloads
    return {"count": workloads.mc_pi(100000)}


def _black_scholes():
    from bohrium_amd import workloads
    s = workloads.black_scholes(4099, steps=2)
    return {"sum0": s[0], "sum1": s[1]}


def _black_scholes_repeat():
    # do_while body with sliding views (the time offset and the output slot move per repeat)
    from bohrium_amd import workloads
    S, K, T = workloads.black_scholes_inputs(4099)
    return {"sums": workloads.black_scholes_repeat(S, K, T, 3)}


def _jacobi():
    from bohrium_amd import workloads
    G = workloads.jacobi_grid(64)
    d = workloads.jacobi_repeat(G, 3)
    return {"grid": G, "delta": d}


def _host_data():
    # arrays from host data, gather, scatter, axis reductions and an int64 scan
    import bohrium_amd as bh
    rng = np.random.RandomState(3)
    a = bh.array(rng.uniform(-1, 1, (20, 30)))
    idx = bh.array(rng.randint(0, 600, 200).astype(np.int64))
    g = bh.take(a, idx)
    t = bh.array(np.zeros(600))
    bh.put(t, bh.array(rng.permutation(600)[:100].astype(np.int64)), bh.array(rng.uniform(size=100)))
    rows = a.sum(axis=1)
    cols = bh.ufunc("maximum").reduce(a, 0)
    k = (a > 0.0).astype(np.int64).cumsum(axis=1)
    return {"gathered": g, "scattered": t, "rows": rows, "cols": cols, "scan": k}


def _batched():
    from bohrium_amd import _bh_api as api, workloads
    sizes = workloads.batched_sizes(16, lo=256, hi=4096)
    api.set_flush_threshold(10 ** 7)
    xs, idxs = workloads.batched_inputs(sizes)
    sums, cums, gathered = workloads.batched_step(xs, idxs)
    api.set_flush_threshold(1000)
    out = {}
    for i in range(len(sizes)):
        out["sum%d" % i], out["cum%d" % i], out["gather%d" % i] = sums[i], cums[i], gathered[i]
    return out


PROGRAMS = {"mc_pi_1e5": _mc_pi, "black_scholes_4099x2": _black_scholes,
            "black_scholes_repeat_4099x3": _black_scholes_repeat, "jacobi_64x3": _jacobi,
            "host_data_gather_scatter": _host_data, "batched_16": _batched}


def record(name, path):
    """Run program `name` while recording its flushes to `path`; returns {result: (address, dtype)}."""
    import bohrium_amd as bh
    from bohrium_amd import _bh_api as api
    api.record_flushes(path)
    try:
        res = PROGRAMS[name]()
        bh.flush()
        names = {k: (api.base_address(v.base.handle), np.dtype(v.dtype).str) for k, v in res.items()}
    finally:
        api.record_flushes(None)
    return names, res


def main():
    os.makedirs(OUT, exist_ok=True)
    for name in PROGRAMS:
        path = os.path.join(OUT, name + ".bhir")
        names, res = record(name, path)
        with open(os.path.join(OUT, name + ".json"), "w") as f:
            json.dump({"results": {k: {"address": a, "dtype": d} for k, (a, d) in sorted(names.items())}}, f,
                      indent=1, sort_keys=True)
        del res
        print("%s: %d bytes, %d results" % (name, os.path.getsize(path), len(names)))


if __name__ == "__main__":
    main()
