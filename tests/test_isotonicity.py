"""Inclusion isotonicity of the outward-rounded interval operations, on the
oracle's binary64 interval arithmetic (which the device matches bit for bit):
for nested operands C within U, op(C) is possibly undefined only where op(U)
is, and wherever op(U) is defined op(C) lies within it.  (Where op(U) is
possibly undefined its bounds need not contain op(C)'s: the square root of a
negative C is the undefined whole line while that of a U crossing zero is
[0, sqrt(ub)], undefined too.  A comparison with a possibly undefined operand
is never definite, so such values never prune.)  The group tests of the static
levels (a group of consecutive siblings that fails a row on one box holds only
children the per-child test prunes) rest on this property; a bound computed
from a rounded intermediate that is not monotone in its operands would fail
here."""
import numpy as np
import pytest

import pyoracle

N = 40000


def _outer(rng, n):
    """Intervals of mixed signs and scales, some with zero or infinite ends."""
    a = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n).astype(float))
    b = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n).astype(float))
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    k = rng.integers(0, 40, n)
    lo[k == 0] = 0.0
    hi[k == 1] = 0.0
    lo[k == 2] = -np.inf
    hi[k == 3] = np.inf
    lo[k == 4] = -0.0
    hi = np.maximum(lo, hi)
    und = (rng.integers(0, 10, n) == 0).astype(float)
    return np.stack([lo, hi, und], axis=1)


def _inner(rng, u):
    """A random sub-interval of each row of u (ends strictly inside when finite)."""
    lo, hi = u[:, 0].copy(), u[:, 1].copy()
    with np.errstate(invalid="ignore"):
        fl = np.where(np.isfinite(lo), lo, np.where(np.isfinite(hi), hi - np.abs(hi) - 1.0, -1.0))
        fh = np.where(np.isfinite(hi), hi, np.where(np.isfinite(lo), lo + np.abs(lo) + 1.0, 1.0))
    fh = np.maximum(fl, fh)
    t = np.sort(rng.random((len(u), 2)), axis=1)
    a = fl + (fh - fl) * t[:, 0]
    b = fl + (fh - fl) * t[:, 1]
    a = np.clip(a, lo, hi)
    b = np.clip(b, lo, hi)
    keep = rng.integers(0, 4, len(u))            # sometimes keep an end as it is
    a = np.where(keep == 0, lo, a)
    b = np.where(keep == 1, hi, b)
    b = np.maximum(a, b)
    und = u[:, 2] * (rng.integers(0, 2, len(u)))  # undefined only where the outer one may be
    return np.stack([a, b, und], axis=1)


def _check_nested(inner, outer, what):
    # the undefined flag is monotone
    assert not ((inner[:, 2] > 0) & (outer[:, 2] == 0)).any(), what
    # where the outer result is defined, the inner one lies within it
    defined = outer[:, 2] == 0
    assert defined.mean() > 0.2, what
    i, o = inner[defined], outer[defined]
    assert not np.isnan(o[:, :2]).any() and not np.isnan(i[:, :2]).any(), what
    bad = ~((i[:, 0] >= o[:, 0]) & (i[:, 1] <= o[:, 1]))
    assert not bad.any(), (what, i[bad][:3], o[bad][:3])


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "max", "min"])
def test_binary_ops_isotone(op):
    rng = np.random.default_rng(sum(map(ord, op)))
    ua, ub = _outer(rng, N), _outer(rng, N)
    ca, cb = _inner(rng, ua), _inner(rng, ub)
    _check_nested(pyoracle.interval_op(op, ca, cb), pyoracle.interval_op(op, ua, ub), op)


@pytest.mark.parametrize("op", ["sqrt", "square", "neg", "acos"])
def test_unary_ops_isotone(op):
    rng = np.random.default_rng(sum(map(ord, op)) + 1)
    u = _outer(rng, N)
    if op == "acos":                      # arguments near [-1, 1], some reaching past it
        u[:, :2] = np.sort(np.clip(rng.uniform(-1.3, 1.3, (N, 2)), -2, 2), axis=1)
    c = _inner(rng, u)
    _check_nested(pyoracle.interval_op(op, c), pyoracle.interval_op(op, u), op)


def test_narrow_boxes_isotone():
    # the static-grid case: a parent's subrange against one of its children
    rng = np.random.default_rng(9)
    base = rng.uniform(0.1, 2.0, N)
    width = base * np.exp2(-rng.integers(8, 30, N).astype(float))
    u = np.stack([base, base + width, np.zeros(N)], axis=1)
    t = np.sort(rng.random((N, 2)), axis=1)
    c = np.stack([base + width * t[:, 0], base + width * t[:, 1], np.zeros(N)], axis=1)
    for op in ("sqrt", "square", "acos"):
        uu, cc = u.copy(), c.copy()
        if op == "acos":
            uu[:, :2] -= 1.05
            cc[:, :2] -= 1.05
        _check_nested(pyoracle.interval_op(op, cc), pyoracle.interval_op(op, uu), op)
    for op in ("mul", "div", "sub"):
        _check_nested(pyoracle.interval_op(op, c, c[::-1]), pyoracle.interval_op(op, u, u[::-1]), op)
