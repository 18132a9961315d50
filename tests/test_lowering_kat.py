"""The lowering (squares-in-disk_amd/host/ivarp) against the known answers of
the reference's own lowering tests, and against exact rationals.

tests/golden/lowering_kat.json holds the static_assert / REQUIRE cases of
ivarp/test/fixed_point_bounds.cpp, compile_time_bounds.cpp,
bound_and_simplify.cpp, bound_dependency_analysis.cpp and rewrite_to_bounds.cpp,
each with its source line.  tests/lowering_driver.cpp runs the lowering's passes
on them.  The reference's randomized tests of the same files (fixed-point
results bracket the exact rational result and are at most one unit apart;
interval products and quotients contain the exact ones; sqrt bounds bracket
the root) are checked here with Python's exact fractions, independently of the
lowering's own big-rational code.
"""
import json
import random
import shutil
import subprocess
from fractions import Fraction
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "squares-in-disk_amd"
GOLD = json.loads((ROOT / "tests" / "golden" / "lowering_kat.json").read_text())
DEN = 1000000
MAXB = 9223372036854775807
MINB = -MAXB


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path_factory.mktemp("lowering") / "lowering_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-I", str(PKG / "host"), "-I", str(PKG / "host" / "ivarp"),
           str(ROOT / "tests" / "lowering_driver.cpp"), str(PKG / "host/ivarp/expr.cpp"),
           str(PKG / "host/ivarp/lower.cpp"), str(PKG / "host/ivarp/program.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]

    def run(lines):
        out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
        res = out.stdout.splitlines()
        assert len(res) == len(lines)
        for line, o in zip(lines, res):
            assert not o.startswith("ERROR"), (line, o)
        return res
    return run


def _bounds(args):
    return "%d %s" % (len(args), " ".join("%d %d" % (a, b) for a, b in args))


def test_fixed_point_known_answers(driver):
    cases = GOLD["fixed_point"]
    out = driver(["fp %s %s" % (c["fn"], " ".join(str(a) for a in c["args"])) for c in cases])
    for c, o in zip(cases, out):
        want = c["want"]
        got = o if c["fn"] == "iv_order" else [int(x) for x in o.split()]
        if isinstance(want, int):
            got = got[0]
        assert got == want, (c["src"], c["fn"], c["args"], got, want)
    rel = GOLD["fixed_point_relations"]
    out = driver(["fp %s %s" % (c["fn"], " ".join(str(a) for a in c["args"])) for c in rel])
    for c, o in zip(rel, out):
        lb, ub = (int(x) for x in o.split())
        assert lb < c["lb_lt"] and ub > c["ub_gt"], (c["src"], lb, ub)


def test_compile_time_predicate_bounds(driver):
    cases = GOLD["ct_predicate_bounds"]
    out = driver(["ctb %s %s" % (_bounds(c["args"]), c["expr"]) for c in cases])
    for c, o in zip(cases, out):
        kind, plb, pub = o.split()
        assert kind == "P" and (plb == "1", pub == "1") == (c["want"]["plb"], c["want"]["pub"]), (c["src"], o)


def test_bound_and_simplify(driver):
    cases = GOLD["bound_and_simplify"]
    out = driver(["bs %s %s" % (_bounds(c["args"]), c["expr"]) for c in cases])
    for c, o in zip(cases, out):
        kind, a, b = o.split()
        if "lb" in c["want"]:
            assert kind == "E" and (int(a), int(b)) == (c["want"]["lb"], c["want"]["ub"]), (c["src"], o)
        else:
            assert kind == "P" and (a == "1", b == "1") == (c["want"]["plb"], c["want"]["pub"]), (c["src"], o)


def test_bound_dependencies(driver):
    lines, want = [], []
    for c in GOLD["bound_dependencies"]:
        for arg, w in enumerate(c["want"]):
            lines.append("dep %s %d %s" % (_bounds(c["args"] if len(c["args"]) > arg else c["args"] + [[MINB, MAXB]]),
                                            arg, c["expr"]))
            want.append((c["src"], arg, w))
    out = driver(lines)
    for (src, arg, w), o in zip(want, out):
        assert [int(x) for x in o.split()] == w, (src, arg, o, w)


def test_rewrite_to_bounds(driver):
    cases = GOLD["rewrite_to_bounds"]
    out = driver(["rw %d %s" % (c["target"], c["expr"]) for c in cases])
    for c, o in zip(cases, out):
        f = o.split()
        assert int(f[0]) == len(c["want"]), (c["src"], o)
        for k, w in enumerate(c["want"]):
            d, clb, cub, ctl, ctu = f[1 + 5 * k: 6 + 5 * k]
            assert d == w["dir"] and float(clb) == w["value"] and float(cub) == w["value"], (c["src"], o)
            assert [int(ctl), int(ctu)] == w["ct"], (c["src"], o)
    cases = GOLD["successful_rewrite"]
    out = driver(["rw %d %s" % (c["target"], c["expr"]) for c in cases])
    for c, o in zip(cases, out):
        assert (int(o.split()[0]) > 0) == c["want"], (c["src"], o)


def test_constraint_rows(driver):
    for c in GOLD["constraint_rows"]:
        o = driver(["rows %d %d %s" % (c["nargs"], len(c["exprs"]), " ".join(c["exprs"]))])[0]
        rows = [[int(x) for x in part.split()] for part in o.split("|")[1:]]
        assert rows == c["want"], (c["src"], o)


def _rat(v):
    return Fraction(v, DEN)


def test_fixed_point_randomized_against_exact(driver):
    # fixed_point_bounds.cpp:40-72, 90-129, 237-252 and compile_time_bounds.cpp:45-181, checked with fractions
    rng = random.Random(20261203)
    small = [rng.randint(-3000000000, 3000000000) for _ in range(16)]
    big = [rng.randint(-300000000000000, 300000000000000) for _ in range(16)]
    vals = small + big
    lines, meta = [], []
    for i, a in enumerate(vals):
        for j, b in enumerate(vals):
            lines += ["fp mul_rd %d %d" % (a, b), "fp mul_ru %d %d" % (a, b)]
            meta.append(("mul", a, b, i < 16 and j < 16))
            if b != 0:
                lines += ["fp div_rd %d %d" % (a, b), "fp div_ru %d %d" % (a, b)]
                meta.append(("div", a, b, abs(b) >= 1000 and abs(a) <= 5000000000000000))
    out = iter(driver(lines))
    for kind, a, b, finite in meta:
        lo, hi = int(next(out)), int(next(out))
        exact = Fraction(a * b, DEN * DEN) if kind == "mul" else Fraction(a, b)
        assert lo <= hi, (kind, a, b)
        if finite:
            assert lo > MINB and hi < MAXB, (kind, a, b, lo, hi)
        if lo > MINB:
            assert _rat(lo) <= exact, (kind, a, b, lo)
        if hi < MAXB:
            assert exact <= _rat(hi), (kind, a, b, hi)
        if finite and kind == "mul" and abs(a) <= 1000 * DEN and abs(b) <= 1000 * DEN:
            assert hi - lo <= 1, (kind, a, b, lo, hi)
    # add/sub/mul/div of intervals contain the exact results (compile_time_bounds.cpp:64-147)
    lines, meta = [], []
    for _ in range(400):
        x = sorted(rng.randint(-1000 * DEN, 1000 * DEN) for _ in range(2))
        y = sorted(rng.randint(-1000 * DEN, 1000 * DEN) for _ in range(2))
        lines += ["fp add_rd %d %d" % (x[0], y[0]), "fp add_ru %d %d" % (x[1], y[1]),
                  "fp iv_mul_lb %d %d %d %d" % (x[0], x[1], y[0], y[1]),
                  "fp iv_mul_ub %d %d %d %d" % (x[0], x[1], y[0], y[1]),
                  "fp iv_div_lb %d %d %d %d" % (x[0], x[1], y[0], y[1]),
                  "fp iv_div_ub %d %d %d %d" % (x[0], x[1], y[0], y[1])]
        meta.append((x, y))
    out = iter(driver(lines))
    for x, y in meta:
        al, au, ml, mu, dl, du = (int(next(out)) for _ in range(6))
        X, Y = [_rat(v) for v in x], [_rat(v) for v in y]
        assert _rat(al) <= X[0] + Y[0] and X[1] + Y[1] <= _rat(au)
        prods = [p * q for p in X for q in Y]
        assert _rat(ml) <= min(prods) and max(prods) <= _rat(mu)
        if y[0] > 0 or y[1] < 0:
            quots = [p / q for p in X for q in Y]
            assert dl > MINB and du < MAXB and _rat(dl) <= min(quots) and max(quots) <= _rat(du)
        else:
            assert dl == MINB and du == MAXB        # a divisor that may be zero: unbounded
    # sqrt bounds bracket the root; fixed_pow of the sqrt bounds brackets x (fixed_point_bounds.cpp:237-252)
    xs = [rng.randint(0, DEN) for _ in range(200)] + [rng.randint(DEN, MAXB - 1) for _ in range(200)]
    out = iter(driver(sum([["fp sqrt_rd %d" % x, "fp sqrt_ru %d" % x] for x in xs], [])))
    for x in xs:
        rd, ru = int(next(out)), int(next(out))
        assert rd <= ru and _rat(rd) ** 2 <= _rat(x) <= _rat(ru) ** 2, (x, rd, ru)
    # interval order against exact comparisons (fixed_point_bounds.cpp:149-196)
    pts = [rng.randint(-300000000, 300000000) for _ in range(16)]
    lines, meta = [], []
    for a in pts:
        for b in pts:
            if b < a:
                continue
            for c in pts[:8]:
                for d in pts[:8]:
                    if d < c:
                        continue
                    lines.append("fp iv_order %d %d %d %d" % (a, b, c, d))
                    meta.append((a, b, c, d))
    out = driver(lines)
    for (a, b, c, d), o in zip(meta, out):
        if o == "LT":
            assert b < c
        elif o == "LE":
            assert b <= c and not b < c
        elif o == "EQ":
            assert a == b == c == d
        elif o == "GT":
            assert a > d
        elif o == "GE":
            assert a >= d and not a > d
        else:
            assert not (b <= c) and not (a >= d), (a, b, c, d)
