"""Derives the constants of kernel_math.inc: the exp polynomial and the log table.

    python bohrium_amd/csrc/ve_hip/gen_kernel_math.py          # prints the C++ text of both

exp: after the reduction x = n ln2 + r, |r| <= ln2/2, exp(r) = 1 + r + r^2 q(r) where q interpolates
(exp(r) - 1 - r) / r^2 at the 10 Chebyshev nodes of [-a, a] (a slightly above ln2/2, to cover the
rounding of the reduction), evaluated in 60-digit arithmetic; the coefficients are rounded to double.
The script reports the largest relative error of the rounded polynomial against exp over 20001 points.

log (Tang's table method): the frexp mantissa is moved into [0.70703125, 1.4140625); the top 7 bits of
its fraction select one of 128 intervals of width 1/128 (upper binade, j = 0..52) or 1/256 (lower
binade, j = 53..127).  Entry j holds c_j ~ 1/(interval centre) rounded to 9 significant bits (so that
r = m c_j - 1, |r| < 2^-7, takes one fma) and -log(c_j) split into a double head and tail.  The two
intervals next to 1 (j = 0 and j = 127) use c_j = 1, so log(m) = log1p(m - 1) there without cancellation.
Only mpmath is needed, at development time; the engine ships the printed constants."""
import mpmath as mp

mp.mp.dps = 60


def exp_coefficients():
    a = mp.log(2) / 2 * (1 + mp.mpf("1e-6"))

    def g(r):
        if abs(r) < mp.mpf("1e-20"):
            return mp.mpf(1) / 2 + r / 6
        return (mp.exp(r) - 1 - r) / (r * r)

    nodes = [a * mp.cos(mp.pi * (k + mp.mpf(1) / 2) / 10) for k in range(10)]
    # interpolating polynomial of degree 9 through (node, g(node)): solve the Vandermonde system
    A = mp.matrix([[t ** i for i in range(10)] for t in nodes])
    b = mp.matrix([g(t) for t in nodes])
    c = mp.lu_solve(A, b)
    coeffs = [float(c[i]) for i in range(10)]  # q(r) = sum coeffs[i] r^i, i.e. c_{i+2} of exp

    worst = mp.mpf(0)
    for k in range(20001):
        r = -a + 2 * a * k / 20000
        q = mp.mpf(coeffs[9])
        for i in range(8, -1, -1):
            q = q * r + mp.mpf(coeffs[i])
        p = 1 + r + r * r * q
        worst = max(worst, abs(p / mp.exp(r) - 1))
    return coeffs, worst


def round_sig(v, bits):
    """v rounded to `bits` significant bits."""
    e = int(mp.floor(mp.log(v, 2)))
    scale = mp.mpf(2) ** (bits - 1 - e)
    return mp.nint(v * scale) / scale


def log_table():
    rows = []
    for j in range(128):
        if j <= 52:
            centre = 1 + (mp.mpf(j) + mp.mpf(1) / 2) / 128
        else:
            centre = (1 + (mp.mpf(j) + mp.mpf(1) / 2) / 128) / 2
        c = mp.mpf(1) if j in (0, 127) else round_sig(1 / centre, 9)
        L = -mp.log(c)
        hi = float(L)
        lo = float(L - mp.mpf(hi))
        rows.append((float(c), hi, lo))
    return rows


def max_reduced_argument(rows):
    """The largest |m c_j - 1| over each interval's ends."""
    worst = mp.mpf(0)
    for j, (c, _, _) in enumerate(rows):
        if j <= 52:
            lo, hi = 1 + mp.mpf(j) / 128, 1 + mp.mpf(j + 1) / 128
        else:
            lo, hi = (1 + mp.mpf(j) / 128) / 2, (1 + mp.mpf(j + 1) / 128) / 2
        if j == 53:
            lo = mp.mpf("0.70703125")
        if j == 52:
            hi = mp.mpf("1.4140625")
        worst = max(worst, abs(lo * c - 1), abs(hi * c - 1))
    return worst


def text():
    coeffs, err = exp_coefficients()
    rows = log_table()
    out = ["// exp: q(r) coefficients c2..c11 (max relative error of 1 + r + r^2 q(r): %.2e)" % float(err)]
    for i in range(10):
        out.append("    const double c%d = %s;" % (i + 2, float.hex(coeffs[i])))
    out.append("// log: %d entries {c_j, -log(c_j) head, tail, 0}; max |m c_j - 1| = %.5f (2^-7 = %.5f)"
               % (len(rows), float(max_reduced_argument(rows)), 2.0 ** -7))
    out.append("BH_MATH_TABLE double bh_log_tab[128][4] = {")
    for c, hi, lo in rows:
        out.append("    {%s, %s, %s, 0.0}," % (float.hex(c), float.hex(hi), float.hex(lo)))
    out.append("};")
    return "\n".join(out)


if __name__ == "__main__":
    print(text())
