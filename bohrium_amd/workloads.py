"""The benchmark programs, written against the bh107-style front end.

They are ordinary array programs; the engine sees only their bytecode.
    mc_pi          Monte Carlo pi: philox uniforms x, y; count(x*x + y*y <= 1) as int64
    black_scholes  fused option pricing with the Abramowitz-Stegun CND polynomial, one sum per step
    jacobi         5-point heat stencil on offset views with a max-abs-delta per iteration
"""
import numpy as np

from . import bharray as _b
from .bharray import BhArray, ufunc
from .random import RandomState

# Abramowitz & Stegun 26.2.17 coefficients of the normal CDF polynomial.
CND_A = (0.31938153, -0.356563782, 1.781477937, -1.821255978, 1.330274429)
CND_P = 0.2316419
RSQRT2PI = 0.39894228040143267794


def mc_pi(n, seed=42):
    """x, y = philox uniforms (counters 0..n-1 and n..2n-1, key `seed`); returns the int64 count."""
    R = RandomState(seed)
    x = R.random_sample(n)
    y = R.random_sample(n)
    return ((x * x + y * y) <= 1.0).astype(np.int64).sum()


def black_scholes_inputs(n, seed=42):
    R = RandomState(seed)
    S = R.uniform(58.0, 62.0, n)
    K = R.uniform(55.0, 65.0, n)
    T = R.uniform(0.25, 2.0, n)
    return S, K, T


def _cnd(d):
    L = ufunc("absolute")(d)
    k = 1.0 / (L * CND_P + 1.0)
    a1, a2, a3, a4, a5 = CND_A
    poly = k * (a1 + k * (a2 + k * (a3 + k * (a4 + k * a5))))
    w = 1.0 - RSQRT2PI * ufunc("exp")(L * L * -0.5) * poly
    neg = d < 0.0
    return w + neg.astype(np.float64) * (1.0 - 2.0 * w)


def black_scholes_step(S, K, T, step, r=0.02, sigma=0.3, dt=0.01):
    """Sum of European call prices at time-to-expiry T + step*dt (one fused map + reduce)."""
    Tt = T + step * dt
    sq = ufunc("sqrt")(Tt)
    d1 = (ufunc("log")(S / K) + (r + 0.5 * sigma * sigma) * Tt) / (sigma * sq)
    d2 = d1 - sigma * sq
    call = S * _cnd(d1) - K * ufunc("exp")(Tt * -r) * _cnd(d2)
    return call.sum()


def black_scholes(n, steps=10, seed=42):
    S, K, T = black_scholes_inputs(n, seed)
    return [black_scholes_step(S, K, T, i) for i in range(steps)]


def black_scholes_repeat(S, K, T, steps=10, r=0.02, sigma=0.3, dt=0.01):
    """The same `steps` time steps as one recorded loop body replayed by the engine
    (do_while): step i reads the time offset dts[i] through a sliding view and writes
    its sum to sums[i].  Returns the float64 array of the `steps` sums."""
    dts = _b.arange(steps, dtype=np.float64) * dt
    sums = _b.empty(steps, np.float64)

    def body():
        i = _b.get_iterator()
        Tt = T + dts[i]
        sq = ufunc("sqrt")(Tt)
        d1 = (ufunc("log")(S / K) + (r + 0.5 * sigma * sigma) * Tt) / (sigma * sq)
        d2 = d1 - sigma * sq
        call = S * _cnd(d1) - K * ufunc("exp")(Tt * -r) * _cnd(d2)
        sums[i] = call.sum()

    from . import do_while
    do_while(body, steps)
    return sums


def jacobi_grid(n, dtype=np.float32):
    """n x n grid: top row 40.0, the other borders -273.0, interior 0."""
    G = _b.zeros((n, n), dtype)
    G[:, 0] = -273.0
    G[:, -1] = -273.0
    G[-1, :] = -273.0
    G[0, :] = 40.0
    return G


def jacobi_step(G):
    """One Jacobi sweep in place; returns the max |new - old| of the interior (0-d array)."""
    c = G[1:-1, 1:-1]
    tmp = (c + G[:-2, 1:-1] + G[2:, 1:-1] + G[1:-1, :-2] + G[1:-1, 2:]) * 0.2
    delta = ufunc("absolute")(tmp - c).max()
    c[...] = tmp
    return delta


def jacobi(n, iters):
    G = jacobi_grid(n)
    deltas = [jacobi_step(G) for _ in range(iters)]
    return G, deltas


def batched_sizes(count=4096, lo=256, hi=262144, seed=42):
    """Array sizes floor(exp(U[ln lo, ln hi])) from NumPy's seeded generator (host-side sizes)."""
    rng = np.random.RandomState(seed)
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)


def batched_inputs(sizes, seed=42):
    """Per size n: x ~ N(0, 1) float32 of n elements and n int64 indices uniform in [0, n) (philox)."""
    R = RandomState(seed)
    xs = [R.standard_normal(int(n), dtype=np.float32) for n in sizes]
    idxs = [R.randint(0, int(n), int(n)) for n in sizes]
    return xs, idxs


def batched_step(xs, idxs, a=0.5, b=0.1):
    """Per array: y = tanh(a*x + b), sum(y), the int64 cumsum of (x > 0) and y[idx]."""
    sums, cums, gathered = [], [], []
    for x, idx in zip(xs, idxs):
        y = ufunc("tanh")(x * a + b)
        sums.append(y.sum())
        cums.append((x > 0).astype(np.int64).cumsum())
        gathered.append(_b.take(y, idx))
    return sums, cums, gathered


def random_dag(seed, n_instr=100, sides=(512, 1024, 2048), live=4):
    """A seeded random array program of about `n_instr` bytecode instructions on float32
    arrays: element-wise add/multiply/exp/sqrt(|x| + 1), axis reductions (broadcast back), transposed,
    reversed and half-size slice views.  At most `live` arrays per shape stay referenced
    (older ones are freed, so the fuser can turn them into temporaries).  Returns the
    arrays still live at the end; fetching them runs the whole program."""
    rng = np.random.RandomState(seed)
    s = int(rng.choice(sides))
    R = RandomState(seed)
    pools = {}

    def add(a):
        p = pools.setdefault(a.shape, [])
        p.append(a)
        if len(p) > live:
            p.pop(0)

    def pick(shape=None):
        shapes = [k for k, v in pools.items() if v and (shape is None or k == shape)]
        k = shapes[rng.randint(len(shapes))]
        return pools[k][rng.randint(len(pools[k]))]

    for _ in range(3):
        add(R.random_sample((s, s)).astype(np.float32) * 2.0 - 1.0)
    count = 0
    sqrt, exp, absolute = ufunc("sqrt"), ufunc("exp"), ufunc("absolute")
    while count < n_instr:
        op = rng.choice(["add", "mul", "exp", "sqrt", "reduce", "transpose", "reverse", "half"],
                        p=[0.25, 0.2, 0.1, 0.1, 0.1, 0.1, 0.1, 0.05])
        a = pick()
        if op in ("add", "mul"):
            b = pick(a.shape)
            if rng.rand() < 0.3 and a.ndim == 2 and a.shape[0] == a.shape[1]:
                b = b.T
            c = (a + b) * 0.5 if op == "add" else a * b
            count += 2 if op == "add" else 1
        elif op == "exp":
            c = exp(a) * 0.3
            count += 2
        elif op == "sqrt":
            # (+1 keeps the program well conditioned: x - mean(x) of a constant row is rounding
            # noise, and the square root of noise differs by orders of magnitude between engines)
            c = sqrt(absolute(a) + 1.0)
            count += 3
        elif op == "reduce" and a.ndim == 2:
            axis = int(rng.randint(2))
            m = a.sum(axis=axis) * (1.0 / a.shape[axis])
            c = a - (m[None, :] if axis == 0 else m[:, None])
            count += 3
        elif op == "transpose" and a.ndim == 2 and a.shape[0] == a.shape[1]:
            c = a.T * 1.0
            count += 1
        elif op == "reverse" and a.ndim == 2:
            c = a[::-1, :] - a[:, ::-1]
            count += 1
        elif op == "half" and a.ndim == 2 and a.shape[0] >= 4:
            h = a.shape[0] // 2
            c = a[:h] + a[h:2 * h]
            count += 1
        else:
            continue
        add(c)
    return [a for p in pools.values() for a in p]


def jacobi_repeat(G, iters):
    """`iters` Jacobi sweeps of G in place as one recorded body replayed by the engine
    (do_while); returns the 0-d array with the last sweep's max |new - old|."""
    delta = _b.empty((), G.dtype)

    def body():
        delta[...] = jacobi_step(G)

    from . import do_while
    do_while(body, iters)
    return delta


if __name__ == "__main__":
    # Small drivers for profiling runs:
    #   python -m bohrium_amd.workloads jacobi [n] [sweeps]       (default 8192, 5)
    #   python -m bohrium_amd.workloads black_scholes [n] [steps] (default 20000000, 10)
    #   python -m bohrium_amd.workloads mc_pi [n] [runs]          (default 10000000, 40)
    #   python -m bohrium_amd.workloads batched [arrays] [runs]   (default 4096, 5)
    import sys
    what = sys.argv[1] if len(sys.argv) > 1 else "jacobi"
    if what == "jacobi":
        n = int(sys.argv[2]) if len(sys.argv) > 2 else 8192
        it = int(sys.argv[3]) if len(sys.argv) > 3 else 5
        grid = jacobi_grid(n)
        print("jacobi %d x %d, %d sweeps: last max |delta| %r" % (n, n, it, float(jacobi_repeat(grid, it).asnumpy())))
    elif what == "black_scholes":
        n = int(sys.argv[2]) if len(sys.argv) > 2 else 20_000_000
        st = int(sys.argv[3]) if len(sys.argv) > 3 else 10
        S, K, T = black_scholes_inputs(n)
        print("black-scholes n=%d: sums %s" % (n, black_scholes_repeat(S, K, T, st).asnumpy()[:3]))
    elif what == "mc_pi":
        n = int(sys.argv[2]) if len(sys.argv) > 2 else 10_000_000
        reps = int(sys.argv[3]) if len(sys.argv) > 3 else 40
        for _ in range(reps):
            c = int(mc_pi(n).asnumpy())
        print("mc_pi n=%d: count %d" % (n, c))
    elif what == "batched":
        from . import _bh_api
        count = int(sys.argv[2]) if len(sys.argv) > 2 else 4096
        reps = int(sys.argv[3]) if len(sys.argv) > 3 else 5
        sizes = batched_sizes(count)
        _bh_api.set_flush_threshold(10 ** 7)
        xs, idxs = batched_inputs(sizes)
        _b.flush()
        for _ in range(reps):
            sums, cums, gathered = batched_step(xs, idxs)
            _b.asnumpy_many([sums[-1], cums[-1][:1], gathered[-1][:1]])
            del sums, cums, gathered
        print("batched %d arrays (%d elements), %d runs" % (count, int(sizes.sum()), reps))
    else:
        raise SystemExit("usage: python -m bohrium_amd.workloads jacobi|black_scholes|mc_pi|batched [n] [steps]")
