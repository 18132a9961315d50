"""The lemma lowering (host C++, through the C ABI's program functions; no GPU
calls) against the reference's own prover-input dump of every statement:
argument count, split policy, runtime-bound-table size, constraint-row sizes and
initial-queue size (print_prover_input of the reference build,
ivarp/include/ivarp/prover_input/print_prover_input_impl.hpp), plus initial
boxes, program-text round trips and the exported C ABI."""
import math
import re
from fractions import Fraction
from pathlib import Path

import pytest

import pyoracle
import sqd

ROOT = Path(__file__).resolve().parent.parent

# key: (N, [(arg, initial, subs, dynamic)], bounds, row sizes, initial queue)
# split policies: src/*_subcontainer.cpp, src/top_packing.cpp, src/more_subcontainers.cpp
REFERENCE_DUMP = {
    "top1": (1, [(0, 128, 8, 1)], 2, [3], 128),
    "top2": (1, [(0, 128, 8, 1)], 2, [3], 128),
    "one": (3, [(0, 128, 8, 1), (1, 0, 256, 0), (2, 0, 256, 0)], 11, [0, 1, 10], 128),
    "two": (4, [(0, 128, 8, 1), (1, 128, 8, 1), (2, 0, 64, 0), (3, 0, 64, 0)], 16, [0, 0, 1, 12], 16384),
    "three": (5, [(0, 64, 8, 1), (1, 64, 8, 1), (2, 16, 4, 1), (3, 0, 64, 0), (4, 0, 32, 0)], 19,
              [0, 0, 0, 1, 14], 65536),
    "four": (6, [(0, 64, 8, 1), (1, 64, 8, 1), (2, 16, 4, 1), (3, 0, 32, 0), (4, 0, 32, 0), (5, 0, 16, 0)], 16,
             [0, 0, 0, 0, 1, 13], 65536),
    "five": (7, [(0, 128, 4, 1), (1, 16, 4, 1), (2, 16, 4, 1), (3, 0, 16, 0), (4, 0, 16, 0), (5, 0, 16, 0),
                 (6, 0, 4, 0)], 30, [0] * 6 + [21], 32768),
    "six": (8, [(0, 128, 4, 1), (1, 16, 4, 1), (2, 16, 4, 1), (3, 16, 4, 1), (4, 0, 16, 0), (5, 0, 16, 0),
                (6, 0, 16, 0), (7, 0, 4, 0)], 35, [0] * 7 + [24], 524288),
    "seven": (9, [(0, 128, 4, 1), (1, 16, 4, 1), (2, 16, 4, 1), (3, 16, 4, 1), (4, 0, 16, 0), (5, 0, 16, 0),
                  (6, 0, 16, 0), (7, 0, 16, 0), (8, 0, 4, 0)], 40, [0] * 8 + [27], 524288),
    "more1": (5, [(0, 128, 4, 1), (1, 64, 4, 1), (2, 64, 4, 1), (3, 0, 32, 0), (4, 0, 32, 0)], 12,
              [1, 0, 0, 3, 13], 524288),
    "more2": (6, [(0, 128, 4, 1), (1, 64, 4, 1), (2, 64, 4, 1), (3, 0, 32, 0), (4, 0, 32, 0), (5, 0, 16, 0)], 14,
              [0, 0, 0, 1, 0, 11], 524288),
}


def parse(text):
    lines = text.splitlines()
    n = int(next(l for l in lines if l.startswith("nargs ")).split()[1])
    splits = [tuple(int(v) for v in l.split()[1:5]) for l in lines if l.startswith("split ")]
    bounds = sum(1 for l in lines if l.startswith("bound "))
    rows = [0] * n
    for l in lines:
        if l.startswith("row "):
            f = l.split()
            rows[int(f[1])] = int(f[2])
    iqs = int(next(l for l in lines if l.startswith("iqs ")).split()[1])
    names = [l.split(maxsplit=2)[2] for l in lines if l.startswith("arg ")]
    return n, splits, bounds, rows, iqs, names


@pytest.mark.parametrize("key", sorted(REFERENCE_DUMP))
def test_lowering_matches_reference_dump(program_text, key):
    n, splits, bounds, rows, iqs, _ = parse(program_text(key))
    assert (n, splits, bounds, rows, iqs) == REFERENCE_DUMP[key]


S1_UB = float("1.2649110640673517")   # outward sqrt(1.6), reference dump "initial range"


@pytest.mark.parametrize("key", ["top1", "top2", "one", "two", "three", "four", "more1"])
def test_initial_box_reference_values(program_text, key):
    lb, ub = pyoracle.Oracle(program_text(key)).initial_box()
    assert (lb[0], ub[0]) == (float("0.29499999999999998"), S1_UB)
    for a in range(1, len(lb)):
        assert (lb[a], ub[a]) == (0.0, S1_UB)


# Exact-rational interval evaluation with directed rounding (IEEE 754 fixes
# each rounded result), for the lower bound of sigma(s_1) over the s_1 range.
INF = float("inf")


def _rd(fr):
    d = float(fr)
    return math.nextafter(d, -INF) if Fraction(d) > fr else d


def _ru(fr):
    d = float(fr)
    return math.nextafter(d, INF) if Fraction(d) < fr else d


def _sqrt_rd(x):
    s = math.sqrt(x)
    while Fraction(s) ** 2 > Fraction(x):
        s = math.nextafter(s, -INF)
    while Fraction(math.nextafter(s, INF)) ** 2 <= Fraction(x):
        s = math.nextafter(s, INF)
    return s


def _sqrt_ru(x):
    s = _sqrt_rd(x)
    return s if Fraction(s) ** 2 == Fraction(x) else math.nextafter(s, INF)


def sigma_lower_bound(xl, xu):
    """Lower bound of sigma(s_1) = ite(..., 0.25 (-s - 2 T_inv(s) + sqrt(8 - (s - 2 T_inv(s))^2)), ...)
    (src/fn_aux.hpp:37-47) for s_1 in [xl, xu] (0 < xl): the join's lower bound is the first branch's."""
    F = Fraction
    m_lo = _rd(1 - F(_ru(F(1, 4) * F(_ru(F(xu) * F(xu))))))   # 1 - 0.25 s^2, s >= 0
    m_hi = _ru(1 - F(_rd(F(1, 4) * F(_rd(F(xl) * F(xl))))))
    t_lo = _rd(F(_sqrt_rd(m_lo)) - F(xu))                     # T_inv = sqrt(max(.,0)) - s
    t_hi = _ru(F(_sqrt_ru(m_hi)) - F(xl))
    d_lo = _rd(F(xl) - 2 * F(t_hi))                           # s - 2 T_inv (2 T_inv exact)
    d_hi = _ru(F(xu) - 2 * F(t_lo))
    sq_hi = _ru(max(F(d_lo) ** 2, F(d_hi) ** 2))             # square of a sign-mixed interval
    r_lo = _sqrt_rd(_rd(8 - F(sq_hi)))
    a_lo = _rd(-F(xu) - 2 * F(t_hi))
    return _rd(F(1, 4) * F(_rd(F(a_lo) + F(r_lo))))


@pytest.mark.parametrize("key", ["five", "six", "seven"])
def test_initial_box_sigma_bound(program_text, key):
    lb, ub = pyoracle.Oracle(program_text(key)).initial_box()
    assert (lb[0], ub[0]) == (float("0.29499999999999998"), S1_UB)
    s = sigma_lower_bound(lb[0], ub[0])
    assert s == float("-0.23332051592394792")
    for a in range(1, len(lb)):          # s_n >= sigma(s_1), h_i >= s_n
        assert lb[a] == s and ub[a] == S1_UB


@pytest.mark.parametrize("key", [k for k, _ in sqd.lemmas()] if sqd.LIB_PATH.exists() else [])
def test_program_text_round_trip(key):
    p = sqd.Program.from_lemma(key)
    t = p.text()
    q = sqd.Program.from_text(t)
    assert q.text() == t and q.nargs == p.nargs
    assert q.has_kernels == p.has_kernels


def test_program_text_rejects_garbage():
    with pytest.raises(sqd.SqdError):
        sqd.Program.from_text("sqd-program 1\nnargs 2\n")      # truncated
    with pytest.raises(sqd.SqdError):
        sqd.Program.from_text("not a program")


def test_every_lemma_has_kernels():
    have = {k: sqd.Program.from_lemma(k).has_kernels for k, _ in sqd.lemmas()}
    assert all(have.values()) and len(have) == 14


def test_c_abi_exports_every_declared_function():
    header = (ROOT / "include" / "sqd.h").read_text()
    declared = set(re.findall(r"\b(sqd_[a-z_0-9]+)\s*\(", header))
    lib = sqd.lib()
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing
    assert declared == set(sqd.exported_symbols())


def test_initial_box_more2(program_text):
    # h_i <= T_inv(s_1) (src/more_subcontainers.cpp:109-128): the upper bound is
    # RU(RU(sqrt(RU(1 - RD(0.25 RD(0.295^2))))) - 0.295), evaluated exactly here.
    # (The reference dump quotes 0.69406205568710399, one double lower.)
    lb, ub = pyoracle.Oracle(program_text("more2")).initial_box()
    assert (lb[0], ub[0]) == (float("0.29499999999999998"), S1_UB)
    F = Fraction
    m_hi = _ru(1 - F(_rd(F(1, 4) * F(_rd(F(lb[0]) * F(lb[0]))))))
    t_hi = _ru(F(_sqrt_ru(m_hi)) - F(lb[0]))
    assert t_hi == 0.6940620556871041
    for a in range(1, len(lb)):
        assert (lb[a], ub[a]) == (0.0, t_hi)
