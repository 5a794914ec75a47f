"""float64 exp and log of the generated kernels (bh_exp64 / bh_log64, kernel_math.inc).

CPU: the same source built for the host with g++ (every operation in it is an IEEE operation that
rounds identically on host and device) is checked against NumPy's exp/log, which are correctly
rounded to within 1 ulp, over wide sweeps, subnormals and the special values: at most 2 ulp apart.
GPU: the engine's kernels give the host build's results bit for bit, and so NumPy's within 2 ulp."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "bohrium_amd", "csrc", "ve_hip", "kernel_math.inc")

DRIVER = r"""
#include <cmath>
#include <cstdio>
#include <vector>
using std::fma; using std::rint; using std::frexp; using std::ldexp; using std::fmin; using std::fmax;
#define BH_MATH_FN static inline
#define BH_MATH_TABLE static const
%s
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb");
    std::vector<double> x;
    double v;
    while (std::fread(&v, 8, 1, f) == 1) x.push_back(v);
    std::fclose(f);
    FILE *o = std::fopen(argv[2], "wb");
    for (double a : x) { double r = bh_exp64(a); std::fwrite(&r, 8, 1, o); }
    for (double a : x) { double r = bh_log64(a); std::fwrite(&r, 8, 1, o); }
    std::fclose(o);
    return 0;
}
"""


def inputs():
    r = np.random.RandomState(5)
    parts = [
        r.uniform(-745.5, 710.0, 400000),
        r.uniform(-5, 5, 400000),
        r.uniform(-1e-3, 1e-3, 50000),
        np.exp(r.uniform(-700, 700, 400000)),            # log over the normal range
        r.uniform(0.5, 2.0, 400000),                      # log near 1 and the sqrt(1/2) split
        1.0 + r.uniform(-1e-6, 1e-6, 50000),
        np.ldexp(r.uniform(0.5, 1, 20000), r.randint(-1074, -1021, 20000)),  # subnormals
        np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 709.78, 709.79, 710.0, -708.4, -745.1, -745.2,
                  -746.0, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 0.7071067811865476,
                  0.7071067811865475, 2.0, 0.5, np.e]),
    ]
    return np.concatenate(parts).astype(np.float64)


def host_results(x):
    text = open(SRC).read()
    body = re.search(r'R"BHMATH\((.*)\)BHMATH"', text, re.S).group(1)
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "m.cpp"), os.path.join(d, "m")
        open(src, "w").write(DRIVER % body)
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", src, "-o", exe], check=True)
        x.tofile(os.path.join(d, "in.bin"))
        subprocess.run([exe, os.path.join(d, "in.bin"), os.path.join(d, "out.bin")], check=True)
        out = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64)
    return out[:len(x)], out[len(x):]


def ulp_diff(a, b):
    """|a - b| in units in the last place of b; equal specials count 0, unequal ones inf."""
    both_nan = np.isnan(a) & np.isnan(b)
    same = (a == b) | both_nan
    spacing = np.spacing(np.abs(b))
    with np.errstate(invalid="ignore", over="ignore"):
        d = np.abs(a - b) / spacing
    d = np.where(same, 0.0, d)
    return np.where(np.isfinite(d), d, np.inf)


@pytest.fixture(scope="module")
def host():
    x = inputs()
    return x, host_results(x)


def test_host_build_within_2ulp_of_numpy(host):
    x, (e, l) = host
    with np.errstate(all="ignore"):
        ref_e, ref_l = np.exp(x), np.log(x)
    de, dl = ulp_diff(e, ref_e), ulp_diff(l, ref_l)
    # subnormal results carry fewer bits: compare them in absolute units of the smallest subnormal
    sub = np.abs(ref_e) < 2.2250738585072014e-308
    assert np.all(np.abs(e[sub] - ref_e[sub]) <= 5e-324)
    assert de[~sub].max() <= 2.0, "exp: %g ulp at x=%r" % (de[~sub].max(), x[~sub][np.argmax(de[~sub])])
    assert dl.max() <= 2.0, "log: %g ulp at x=%r" % (dl.max(), x[np.argmax(dl)])


def test_special_values(host):
    x, (e, l) = host
    tail = x[-22:]
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(e[-22:], np.exp(tail))
        np.testing.assert_array_equal(l[-22:], np.log(tail))


@pytest.mark.parametrize("old,new,fn", [
    ("    r = fma(-n, ln2_lo, r);\n", "", "exp"),                          # the tail of ln2 dropped
    ("double q = c11;", "double q = 0.0;", "exp"),                            # the r^11 term dropped
    ("q = fma(q, r, c7);", "q = fma(q, r, -c7);", "exp"),                    # a sign error
    ("p = fma(p, r, 1.0 / 5.0);", "p = fma(p, r, 1.0 / 4.0);", "log"),       # a wrong series coefficient
    ("m < 0.70703125", "m < 0.5", "log"),                                    # no move of m into one range
    (">> 45) & 127u", ">> 44) & 127u", "log"),                               # the table index off by a bit
    ("const double r = fma(m, t[0], -1.0);", "const double r = m * t[0] - 1.0;", "log"),  # r rounded twice
])
def test_a_planted_mistake_is_caught(old, new, fn, monkeypatch, tmp_path):
    """The 2-ulp sweep has teeth: each of these edits to the source fails it."""
    text = open(SRC).read()
    assert old in text
    planted = tmp_path / "kernel_math.inc"
    planted.write_text(text.replace(old, new))
    monkeypatch.setattr(__import__(__name__), "SRC", str(planted))
    x = inputs()
    e, l = host_results(x)
    with np.errstate(all="ignore"):
        d = ulp_diff(e, np.exp(x)) if fn == "exp" else ulp_diff(l, np.log(x))
    if fn == "exp":
        with np.errstate(all="ignore"):
            d = d[np.abs(np.exp(x)) >= 2.2250738585072014e-308]
    assert d.max() > 2.0


@pytest.mark.gpu
def test_device_matches_host_bit_for_bit(host):
    import bohrium_amd as bh
    x, (e, l) = host
    X = bh.array(x)
    de = bh.exp(X).asnumpy()
    dl = bh.log(X).asnumpy()
    assert np.array_equal(de.view(np.uint64), e.view(np.uint64)) or np.array_equal(de, e, equal_nan=True)
    bad = ~((de == e) | (np.isnan(de) & np.isnan(e)))
    assert not bad.any(), "exp differs at %r" % x[bad][:5]
    bad = ~((dl == l) | (np.isnan(dl) & np.isnan(l)))
    assert not bad.any(), "log differs at %r" % x[bad][:5]


def test_constants_are_the_generators():
    """The exp coefficients and the log table in the source are the ones gen_kernel_math.py derives
    (interpolation at Chebyshev nodes and -log(c_j) in 60-digit arithmetic), and each table entry
    satisfies what the log relies on: c_j is 1 next to m = 1 and has at most 9 significant bits, and
    head + tail is -log(c_j) to within 2^-100 relative."""
    mp = pytest.importorskip("mpmath")
    import importlib.util
    path = os.path.join(ROOT, "bohrium_amd", "csrc", "ve_hip", "gen_kernel_math.py")
    spec = importlib.util.spec_from_file_location("gen_kernel_math", path)
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    text = open(SRC).read()
    expected = gen.text()
    for line in expected.splitlines():
        if line.startswith("    const double c") or line.startswith("    {"):
            assert line in text, "kernel_math.inc lacks %r" % line
    rows = gen.log_table()
    assert rows[0][0] == 1.0 and rows[127][0] == 1.0 and rows[0][1] == 0.0 and rows[127][1] == 0.0
    mp.mp.dps = 60
    for c, hi, lo in rows:
        mant = int(mp.ldexp(mp.mpf(c), 8 - int(mp.floor(mp.log(c, 2)))))
        assert mp.mpf(mant) == mp.ldexp(mp.mpf(c), 8 - int(mp.floor(mp.log(c, 2)))), "c = %r has > 9 bits" % c
        if c != 1.0:
            assert abs((mp.mpf(hi) + mp.mpf(lo)) / -mp.log(c) - 1) < mp.mpf(2) ** -100
    assert gen.max_reduced_argument(rows) <= mp.mpf(2) ** -7
