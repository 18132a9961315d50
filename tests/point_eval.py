"""Point evaluation of a lowered program's expressions in near-exact arithmetic.

Test infrastructure for the enclosure checks: the device's interval value of an
expression on a box must contain the expression's real value at every point of
the box (the reference's own tests check its interval functions against point
evaluations on random points, ivarp/test/test_util.hpp:113-235,
test/cuda/cuda_test_fn.cuh:29-86).

A point value is carried as a tiny interval [lo, hi] of dyadic rationals
(fractions.Fraction) that contains the real value: + - * and comparisons of
exact values are exact; every result is rounded outward to PREC significant
bits; sqrt and acos are computed in decimal arithmetic with 150 digits and
widened by 2^-(PREC - 8) of their magnitude, far more than their error and far
less than a binary64 unit in the last place (2^-52).  A value is UNDEF where the
real expression is undefined (division by zero, sqrt of a negative, acos outside
[-1, 1]), and UNKNOWN where the enclosure cannot tell (a divisor or a sqrt
argument whose enclosure straddles zero, a comparison of enclosures that
overlap); checks skip unknown values.

Operations follow the real semantics of the program's node kinds
(squares-in-disk_amd/host/ivarp/program.hpp): if_then_else takes the branch its
condition selects, || and && are the logical connectives on truth values.
"""
import decimal
import struct
from fractions import Fraction

PREC = 400
UNDEF = "undef"
UNKNOWN = "unknown"

_D = decimal.Context(prec=150, rounding=decimal.ROUND_HALF_EVEN)


def _double(bits):
    return struct.unpack("<d", struct.pack("<Q", int(bits)))[0]


def _round(x, up):
    """x rounded to PREC significant bits, downward or upward."""
    if x == 0:
        return x
    e = abs(x.numerator).bit_length() - x.denominator.bit_length()
    shift = PREC - e
    scaled = x * (Fraction(2) ** shift)
    q = scaled.numerator // scaled.denominator                  # floor
    if up and q * scaled.denominator != scaled.numerator:
        q += 1
    return Fraction(q) / (Fraction(2) ** shift)


class V:
    """An enclosure [lo, hi] of a real value."""
    __slots__ = ("lo", "hi")

    def __init__(self, lo, hi=None):
        self.lo = lo
        self.hi = lo if hi is None else hi

    def out(self):
        return V(_round(self.lo, False), _round(self.hi, True))


def _dec(x):
    return _D.divide(decimal.Decimal(x.numerator), decimal.Decimal(x.denominator))


def _from_dec(d, widen):
    """A dyadic enclosure of the decimal value d widened by `widen` relative."""
    f = Fraction(d)
    m = abs(f) * Fraction(1, 2 ** (PREC - 8)) + Fraction(1, 2 ** (PREC + 200)) if widen else Fraction(0)
    return V(f - m, f + m).out()


_PI = None


def _pi():
    global _PI
    if _PI is None:                                            # Machin: pi = 16 atan(1/5) - 4 atan(1/239)
        _PI = _D.subtract(_D.multiply(16, _atan_small(_D.divide(1, 5))), _D.multiply(4, _atan_small(_D.divide(1, 239))))
    return _PI


def _atan_small(t):
    """atan(t) for |t| <= 1/5 by its Taylor series."""
    s = decimal.Decimal(0)
    term = t
    t2 = _D.multiply(t, t)
    k = 0
    eps = decimal.Decimal(10) ** -148
    while abs(term) > eps:
        s = _D.add(s, _D.divide(term, 2 * k + 1) if k % 2 == 0 else -_D.divide(term, 2 * k + 1))
        term = _D.multiply(term, t2)
        k += 1
    return s


def _atan(t):
    """atan(t) for t >= 0."""
    if t > 1:
        return _D.subtract(_D.divide(_pi(), 2), _atan(_D.divide(1, t)))
    halvings = 0
    while t > decimal.Decimal("0.2"):                          # atan t = 2 atan(t / (1 + sqrt(1 + t^2)))
        t = _D.divide(t, _D.add(1, _D.sqrt(_D.add(1, _D.multiply(t, t)))))
        halvings += 1
    return _D.multiply(_atan_small(t), 2 ** halvings)


def _acos_dec(x):
    """acos(x) for -1 <= x <= 1 (decimal)."""
    if x == 1:
        return decimal.Decimal(0)
    if x == -1:
        return _pi()
    # acos x = 2 atan(sqrt((1 - x) / (1 + x)))
    return _D.multiply(2, _atan(_D.sqrt(_D.divide(_D.subtract(1, x), _D.add(1, x)))))


def _sqrt(a):
    if a.hi < 0:
        return UNDEF
    if a.lo < 0:
        return UNDEF if a.hi == 0 and a.lo == 0 else (V(Fraction(0)) if a.hi == 0 else UNKNOWN)
    if a.lo == a.hi:
        x = a.lo
        # exact when x is the square of a dyadic rational
        n, d = x.numerator, x.denominator
        import math
        rn, rd = math.isqrt(n), math.isqrt(d)
        if rn * rn == n and rd * rd == d:
            return V(Fraction(rn, rd))
    lo = _from_dec(_D.sqrt(_dec(a.lo)), True).lo if a.lo > 0 else Fraction(0)
    hi = _from_dec(_D.sqrt(_dec(a.hi)), True).hi
    return V(lo, hi)


def _acos(a):
    one = Fraction(1)
    if a.lo > one or a.hi < -one:
        return UNDEF
    if a.hi > one or a.lo < -one:
        return UNDEF if (a.lo > one or a.hi < -one) else UNKNOWN
    return V(_from_dec(_acos_dec(_dec(a.hi)), True).lo, _from_dec(_acos_dec(_dec(a.lo)), True).hi)


def _mul(a, b):
    ps = [a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi]
    return V(min(ps), max(ps)).out()


def _div(a, b):
    if b.lo == 0 and b.hi == 0:
        return UNDEF
    if b.lo <= 0 <= b.hi:
        return UNKNOWN
    qs = [a.lo / b.lo, a.lo / b.hi, a.hi / b.lo, a.hi / b.hi]
    return V(min(qs), max(qs)).out()


def _cmp(a, b, op):
    """Truth of a op b: True, False or UNKNOWN."""
    if a.lo == a.hi and b.lo == b.hi:                         # exact values
        x, y = a.lo, b.lo
        return {"lt": x < y, "le": x <= y, "gt": x > y, "ge": x >= y, "eq": x == y, "neq": x != y}[op]
    if a.hi < b.lo:
        return {"lt": True, "le": True, "gt": False, "ge": False, "eq": False, "neq": True}[op]
    if a.lo > b.hi:
        return {"lt": False, "le": False, "gt": True, "ge": True, "eq": False, "neq": True}[op]
    return UNKNOWN


class Program:
    """The expression DAG, bounds and constraint rows of a lowered program text."""

    def __init__(self, text):
        self.nodes = {}
        self.bounds = []
        self.rows = {}
        for line in text.splitlines():
            f = line.split()
            if not f:
                continue
            if f[0] == "nargs":
                self.nargs = int(f[1])
            elif f[0] == "node":
                i, op, arg, nk = int(f[1]), f[2], int(f[3]), int(f[4])
                kids = [int(x) for x in f[5:5 + nk]]
                rest = f[5 + nk:]
                node = {"op": op, "arg": arg, "kids": kids, "c_lb": _double(rest[0]), "c_ub": _double(rest[1]),
                        "c_undef": rest[2] == "1", "b_def": rest[3] == "1", "b_poss": rest[4] == "1"}
                self.nodes[i] = node
            elif f[0] == "bound":
                self.bounds.append({"target": int(f[2]), "dir": int(f[3]), "expr": int(f[4])})
            elif f[0] == "row":
                self.rows[int(f[1])] = [int(x) for x in f[3:3 + int(f[2])]]
        roots = set()
        for r in self.rows.values():
            roots.update(r)
        self.constraints = sorted(roots)                       # the device's constraint order (cons_id)

    def evaluate(self, point, roots):
        """Values of `roots` at `point` (a sequence of Fractions): V, True/False,
        UNDEF or UNKNOWN per root."""
        memo = {}

        def ev(i):
            if i in memo:
                return memo[i]
            n = self.nodes[i]
            op, k = n["op"], n["kids"]
            if op == "const":
                r = UNDEF if n["c_undef"] else V(Fraction(n["c_lb"]), Fraction(n["c_ub"]))
            elif op == "arg":
                r = V(point[n["arg"]])
            elif op == "bconst":
                r = True if n["b_def"] else (False if not n["b_poss"] else UNKNOWN)
            elif op == "ite":
                c = ev(k[0])
                r = c if c in (UNDEF, UNKNOWN) else (ev(k[1]) if c else ev(k[2]))
            elif op in ("or", "and"):
                vals = []
                r = None
                for x in k:
                    v = ev(x)
                    if v is (op == "or"):                      # True for ||, False for &&
                        r = v
                        break
                    vals.append(v)
                if r is None:
                    r = UNDEF if UNDEF in vals else (UNKNOWN if UNKNOWN in vals else (op == "and"))
            else:
                args = [ev(x) for x in k]
                if any(a is UNDEF for a in args):
                    r = UNDEF
                elif any(a is UNKNOWN for a in args):
                    r = UNKNOWN
                elif op == "not":
                    r = not args[0]
                elif op in ("lt", "le", "gt", "ge", "eq", "neq"):
                    r = _cmp(args[0], args[1], op)
                else:
                    a = args[0]
                    b = args[1] if len(args) > 1 else None
                    if op == "neg":
                        r = V(-a.hi, -a.lo)
                    elif op == "add":
                        r = V(a.lo + b.lo, a.hi + b.hi).out()
                    elif op == "sub":
                        r = V(a.lo - b.hi, a.hi - b.lo).out()
                    elif op == "mul":
                        r = _mul(a, b)
                    elif op == "div":
                        r = _div(a, b)
                    elif op == "square":
                        sq = [a.lo * a.lo, a.hi * a.hi]
                        r = V(Fraction(0) if a.lo <= 0 <= a.hi else min(sq), max(sq)).out()
                    elif op == "sqrt":
                        r = _sqrt(a)
                    elif op == "acos":
                        r = _acos(a)
                    elif op == "max":                          # n-ary
                        r = V(max(x.lo for x in args), max(x.hi for x in args))
                    elif op == "min":
                        r = V(min(x.lo for x in args), min(x.hi for x in args))
                    else:
                        raise ValueError("unknown node kind " + op)
            memo[i] = r
            return r

        return [ev(r) for r in roots]
