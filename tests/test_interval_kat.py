"""Interval arithmetic of the CPU oracle against the reference's known-answer
interval tests (tests/golden/interval_kat.json, transcribed from
ivarp/test/interval_{addsub,mul,div,sqrt}.cpp and splitter.cpp), against exact
rational arithmetic (directed rounding of + - * / sqrt is fixed by IEEE 754),
and against random rational points in the style of the reference's
random_test_binary_operator (ivarp/test/test_util.hpp:114-198)."""
import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import pyoracle

GOLDEN = json.loads((Path(__file__).parent / "golden" / "interval_kat.json").read_text())
INF = float("inf")


def dec(v):
    return INF if v == "inf" else (-INF if v == "-inf" else float.fromhex(v))


def iv(t):
    return [dec(t[0]), dec(t[1]), float(t[2])]


@pytest.mark.parametrize("case", GOLDEN["cases"], ids=lambda c: c["op"] + "@" + c["src"].split("/")[-1])
def test_oracle_interval_known_answers(case):
    got = pyoracle.interval_op(case["op"], [iv(case["a"])], [iv(case["b"])])[0]
    exp = iv(case["expect"])
    assert got[0] == exp[0] and got[1] == exp[1] and bool(got[2]) == bool(exp[2]), (got, exp)


def rd(fr):
    """Largest double <= the rational fr (finite range)."""
    d = float(fr)
    return math.nextafter(d, -INF) if Fraction(d) > fr else d


def ru(fr):
    d = float(fr)
    return math.nextafter(d, INF) if Fraction(d) < fr else d


def rd_sqrt_exact(x):
    """Largest double s with s*s <= x (x >= 0 a double)."""
    s = math.sqrt(x)
    fx = Fraction(x)
    while Fraction(s) ** 2 > fx:
        s = math.nextafter(s, -INF)
    while Fraction(math.nextafter(s, INF)) ** 2 <= fx:
        s = math.nextafter(s, INF)
    return s


def _random_doubles(rng, n):
    m = rng.uniform(0.5, 1.0, n) * np.exp2(rng.integers(-60, 60, n).astype(float))
    return m * rng.choice([-1.0, 1.0], n)


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "sqrt"])
def test_oracle_directed_rounding_exact(op):
    rng = np.random.default_rng(11 + len(op))
    a = _random_doubles(rng, 3000)
    b = _random_doubles(rng, 3000)
    if op == "sqrt":
        a = np.abs(a)
    lo, hi = pyoracle.rounded_op(op, a, b)
    for x, y, l, h in zip(a, b, lo, hi):
        if op == "sqrt":
            s = rd_sqrt_exact(float(x))
            exp_l, exp_h = s, (s if Fraction(s) ** 2 == Fraction(float(x)) else math.nextafter(s, INF))
        else:
            fx, fy = Fraction(float(x)), Fraction(float(y))
            e = {"add": fx + fy, "sub": fx - fy, "mul": fx * fy, "div": fx / fy}[op]
            exp_l, exp_h = rd(e), ru(e)
        assert (l, h) == (exp_l, exp_h), (op, x, y)


def random_interval(rng, rng_lb, rng_ub, widths):
    wl, wu = widths[rng.integers(len(widths))]
    w = rng.uniform(wl, wu)
    lb = rng.uniform(rng_lb, max(rng_lb, rng_ub - w))
    return [lb, lb + w, 0.0]


def random_point(rng, lb, ub):
    t = Fraction(int(rng.integers(0, 1 << 30)), 1 << 30)
    return Fraction(lb) + (Fraction(ub) - Fraction(lb)) * t


def enclosing(fr):
    return [rd(fr), ru(fr), 0.0]


WIDTHS = [(0, 0.1), (0, 0.5), (0.1, 0.3), (0.5, 2.5), (0.5, 5.0), (5.0, 20.0), (20.0, 50.0)]
RANGES = {  # ivarp/test/interval_mul.cpp:74-78, interval_div.cpp:69-80, interval_addsub.cpp:88
    "add": ([(-100, 100)], [(-100, 100)]),
    "sub": ([(-100, 100)], [(-100, 100)]),
    "mul": ([(-100, 100), (-1000, -900), (900, 1000)], [(-100, 100), (-1000, -900), (900, 1000)]),
    "div": ([(-1000, -900), (900, 1000), (-0.9, -0.2), (0.2, 0.9), (-1000, 1000), (-100, 100), (-0.2, 0.2)],
            [(-1000, -900), (900, 1000), (-0.9, -0.2), (0.2, 0.9)]),
}


def exact(op, p, q):
    return {"add": p + q, "sub": p - q, "mul": p * q, "div": p / q}[op]


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_oracle_interval_random_enclosure(op):
    rng = np.random.default_rng(20261202 + len(op))
    r1, r2 = RANGES[op]
    A, B = [], []
    for _ in range(400):
        A.append(random_interval(rng, *r1[rng.integers(len(r1))], WIDTHS))
        B.append(random_interval(rng, *r2[rng.integers(len(r2))], WIDTHS))
    # pairs of random points and their 3/11, 5/7 scaled versions on 1-ulp intervals
    P = [(random_point(rng, a[0], a[1]), random_point(rng, b[0], b[1])) for a, b in zip(A, B)]
    Ac = [enclosing(p * 3 / 11) for p, _ in P]
    Bc = [enclosing(q * 5 / 7) for _, q in P]
    out = pyoracle.interval_op(op, A, B)
    outc = pyoracle.interval_op(op, Ac, Bc)
    for (p, q), r, rc in zip(P, out, outc):
        if op == "div" and q == 0:
            continue
        e = exact(op, p, q)
        assert (r[0] == -INF or Fraction(r[0]) <= e) and (r[1] == INF or e <= Fraction(r[1]))
        ec = exact(op, p * 3 / 11, q * 5 / 7)
        assert (rc[0] == -INF or Fraction(rc[0]) <= ec) and (rc[1] == INF or ec <= Fraction(rc[1]))
        assert r[0] <= r[1]


def test_oracle_interval_sqrt_points_and_random():
    # ivarp/test/interval_sqrt.cpp:53-69: point intervals of non-squares give adjacent bounds
    for n in GOLDEN["sqrt_points"]["values"]:
        r = pyoracle.interval_op("sqrt", [[n, n, 0.0]])[0]
        assert r[0] < r[1] and math.nextafter(r[0], INF) == r[1] and r[2] == 0.0
        sq = pyoracle.interval_op("square", [r])[0]
        assert sq[0] < n < sq[1]
        assert Fraction(r[0]) ** 2 < n < Fraction(r[1]) ** 2
    # interval_sqrt.cpp:72-92: random intervals in [0, 100]
    rng = np.random.default_rng(5)
    widths = [(0.0, 0.001), (0.0, 0.1), (0.1, 0.5), (0.4, 0.8), (1.0, 2.0), (2.0, 50.0)]
    X = [random_interval(rng, 0, 100, widths) for _ in range(2000)]
    R = pyoracle.interval_op("sqrt", X)
    S = pyoracle.interval_op("square", R)
    for x, r, s in zip(X, R, S):
        assert r[2] == 0.0 and r[0] >= 0
        assert s[0] <= x[0] and s[1] >= x[1]
        assert Fraction(r[0]) ** 2 <= Fraction(x[0]) and Fraction(r[1]) ** 2 >= Fraction(x[1])


def test_oracle_splitter():
    # ivarp/test/splitter.cpp:27-52, and the round-down formula w = RD(RD(ub - lb) / n),
    # p_i = RD(lb + RD(i * w)), p_n = ub (ivarp/include/ivarp/splitter.hpp:44-59)
    n = GOLDEN["splitter"]["n"]
    lo, hi = GOLDEN["splitter"]["range"]
    a = [[lo, hi, 0.0]] * (n + 1)
    b = [[i, n, 0.0] for i in range(n + 1)]
    pts = pyoracle.interval_op("split", a, b)[:, 0]
    assert pts[0] == 0 and pts[n - 1] < 11 and pts[n] == 11
    assert np.all(np.diff(pts) > 0)
    w = rd(Fraction(rd(Fraction(hi) - Fraction(lo))) / n)
    for i in range(n):
        assert pts[i] == rd(Fraction(lo) + Fraction(rd(Fraction(i) * Fraction(w))))
    # an interval whose width is not representable: the grid still tiles it exactly
    a = [[0.1, 1.7, 0.0]] * 9
    pts = pyoracle.interval_op("split", a, [[i, 8, 0.0] for i in range(9)])[:, 0]
    assert pts[0] == 0.1 and pts[8] == 1.7 and np.all(np.diff(pts) > 0)


def _pi_fraction(bits=200):
    """pi as a fraction accurate to 2^-190 (Machin: pi = 16 atan(1/5) - 4 atan(1/239))."""
    def atan_inv(n):
        one = 1 << bits
        s = t = one // n
        k, sign = 1, -1
        while t:
            t //= n * n
            s += sign * (t // (2 * k + 1))
            sign, k = -sign, k + 1
        return s
    return Fraction(16 * atan_inv(5) - 4 * atan_inv(239), 1 << bits)


def test_oracle_acos_known_values():
    # acos as in ivarp/src/interval_acos.cpp:29-66: correctly rounded down / up
    pi = _pi_fraction()
    eps = Fraction(1, 1 << 150)
    cases = [(0.0, pi / 2), (-1.0, pi), (0.5, pi / 3), (-0.5, 2 * pi / 3)]
    for x, v in cases:
        lo, hi = pyoracle.rounded_op("acos", [x], [0.0])
        lo, hi = float(lo[0]), float(hi[0])
        assert Fraction(lo) < v - eps and v + eps < Fraction(hi) and math.nextafter(lo, INF) == hi, x
    lo, hi = pyoracle.rounded_op("acos", [1.0], [0.0])
    assert lo[0] == 0.0 and hi[0] == 0.0
    lo, hi = pyoracle.rounded_op("acos", [1.5, -1.0000000000000002], [0.0, 0.0])
    assert np.all(np.isnan(lo)) and np.all(np.isnan(hi))


def test_oracle_acos_monotone_adjacent():
    rng = np.random.default_rng(9)
    x = np.sort(rng.uniform(-1, 1, 20000))
    lo, hi = pyoracle.rounded_op("acos", x, x)
    assert np.all(np.diff(lo) <= 0) and np.all(np.diff(hi) <= 0)
    up = np.nextafter(lo, INF)
    assert np.array_equal(up, hi)                  # acos of these is irrational: RU is RD's successor
    ref = np.arccos(x)                              # the library result lies within one ulp
    assert np.all(lo <= np.nextafter(ref, INF)) and np.all(hi >= np.nextafter(ref, -INF))


def test_oracle_acos_interval_domain():
    # fn_domain_handle_bounds (math_fn/interval_asin.hpp:13-27) with unbounded compile-time bounds
    r = pyoracle.interval_op("acos", [[-2.0, 0.5, 0.0], [0.1, 0.2, 0.0]])
    lo05, _ = pyoracle.rounded_op("acos", [0.5], [0.0])
    _, hipi = pyoracle.rounded_op("acos", [-1.0], [0.0])
    assert (r[0][0], r[0][1], r[0][2]) == (lo05[0], hipi[0], 1.0)
    assert r[1][2] == 0.0 and r[1][0] < r[1][1]


# ---------------------------------------------------------------- binary32 (FP32 prefilter)
def _rd32(fr):
    """Largest float32 <= the rational fr (finite range)."""
    c = np.float32(float(fr))
    while Fraction(float(c)) > fr:
        c = np.nextafter(c, np.float32(-np.inf))
    while Fraction(float(np.nextafter(c, np.float32(np.inf)))) <= fr:
        c = np.nextafter(c, np.float32(np.inf))
    return float(c)


def _ru32(fr):
    c = np.float32(float(fr))
    while Fraction(float(c)) < fr:
        c = np.nextafter(c, np.float32(np.inf))
    while Fraction(float(np.nextafter(c, np.float32(-np.inf)))) >= fr:
        c = np.nextafter(c, np.float32(-np.inf))
    return float(c)


def _random_floats(rng, n):
    m = rng.uniform(0.5, 1.0, n) * np.exp2(rng.integers(-40, 40, n).astype(float))
    return (m * rng.choice([-1.0, 1.0], n)).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "sqrt"])
def test_oracle_binary32_directed_rounding_exact(op):
    rng = np.random.default_rng(31 + len(op))
    a = _random_floats(rng, 2000)
    b = _random_floats(rng, 2000)
    if op == "sqrt":
        a = np.abs(a)
    lo, hi = pyoracle.rounded_op(op + "32", a, b)
    for x, y, l, h in zip(a, b, lo, hi):
        fx, fy = Fraction(float(x)), Fraction(float(y))
        if op == "sqrt":
            s = _rd32(Fraction(math.sqrt(float(x))))          # candidate, then exact check
            while Fraction(s) ** 2 > fx:
                s = float(np.nextafter(np.float32(s), np.float32(-np.inf)))
            while Fraction(float(np.nextafter(np.float32(s), np.float32(np.inf)))) ** 2 <= fx:
                s = float(np.nextafter(np.float32(s), np.float32(np.inf)))
            exp_l = s
            exp_h = s if Fraction(s) ** 2 == fx else float(np.nextafter(np.float32(s), np.float32(np.inf)))
        else:
            e = {"add": fx + fy, "sub": fx - fy, "mul": fx * fy, "div": fx / fy}[op]
            exp_l, exp_h = _rd32(e), _ru32(e)
        assert (l, h) == (exp_l, exp_h), (op, x, y)


def test_oracle_binary32_outward_conversion_and_acos():
    rng = np.random.default_rng(77)
    x = np.concatenate([rng.uniform(-2, 2, 3000), [0.1, -0.1, 1.0 / 3, 1e-40, -1e-40, 3.5e38, -3.5e38]])
    lo, hi = pyoracle.rounded_op("to32", x, x)
    fmax = float(np.finfo(np.float32).max)
    for v, l, h in zip(x[:-2], lo[:-2], hi[:-2]):
        fr = Fraction(float(v))
        assert (l, h) == (_rd32(fr), _ru32(fr)), v
    assert (lo[-2], hi[-2]) == (fmax, INF) and (lo[-1], hi[-1]) == (-INF, -fmax)   # beyond the float range
    # binary32 acos = binary64 directed results converted outward: adjacent floats around acos(x)
    xs = rng.uniform(-1, 1, 2000).astype(np.float32).astype(np.float64)
    lo32, hi32 = pyoracle.rounded_op("acos32", xs, xs)
    lo64, hi64 = pyoracle.rounded_op("acos", xs, xs)
    assert np.all(lo32 <= lo64) and np.all(hi32 >= hi64)
    assert np.array_equal(np.nextafter(lo32.astype(np.float32), np.float32(np.inf)).astype(np.float64), hi32)
