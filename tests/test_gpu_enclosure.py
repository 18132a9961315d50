"""Lemma-level enclosure: the device's interval of every bound expression and
the IBool of every constraint, on boxes the proofs visit, against the real
values at sampled points of those boxes.

The reference's tests evaluate its interval functions on random points and
intervals and require the interval results to contain the point results
(ivarp/test/test_util.hpp:113-235, test/cuda/cuda_test_fn.cuh:29-86).  Here the
functions are the generated code of each statement (including its compile-time
sign shortcuts, lazy if_then_else and domain clamping), evaluated on the device
in binary64 and, on boxes converted outward, in binary32 (the prefilter), and
the points are evaluated by tests/point_eval.py in near-exact arithmetic:

  bound k:       device [lb, ub] contains the real value wherever it is defined,
                 and is flagged possibly undefined wherever it is not;
  constraint j:  a point where it holds has `possibly`, one where it fails has
                 not `definitely`.

Boxes: the initial box, initial-queue boxes, and random cells of their static
grids and dynamic subdivisions (the boxes the static and dynamic levels see).
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import point_eval
import sqd

pytestmark = pytest.mark.gpu

KEYS = ["top1", "top2", "one", "two", "three", "four", "five", "six", "seven", "more1", "more2", "ec1", "ec2"]


def _splits(text):
    out = []
    for line in text.splitlines():
        f = line.split()
        if f and f[0] == "split":
            out.append((int(f[1]), int(f[2]), int(f[3]), f[4] == "1"))
    return out


def _boxes(core, text, rng, n_queue=12):
    """SoA boxes (N, n): the initial box, initial-queue boxes and random cells of
    their static grids and dynamic subdivisions."""
    blb, bub = core.initial_box()
    qlb, qub, _ = core.initial_queue()
    pick = rng.sample(range(qlb.shape[1]), min(n_queue, qlb.shape[1]))
    lbs = [blb[:, None], qlb[:, pick]]
    ubs = [bub[:, None], qub[:, pick]]
    for dynamic in (False, True):                  # one random cell of every static (dynamic) variable
        clb, cub = qlb[:, pick].copy(), qub[:, pick].copy()
        for arg, _, subs, dyn in _splits(text):
            if dyn != dynamic:
                continue
            for i in range(clb.shape[1]):
                w = (cub[arg, i] - clb[arg, i]) / subs
                c = rng.randrange(subs)
                lo = clb[arg, i] + c * w
                hi = clb[arg, i] + (c + 1) * w if c + 1 < subs else cub[arg, i]
                clb[arg, i], cub[arg, i] = lo, max(lo, min(hi, cub[arg, i]))
        lbs.append(clb)
        ubs.append(cub)
    return np.hstack(lbs), np.hstack(ubs)


def _points(lb, ub, rng, n_random=3):
    """Points of a box: its two corners and random points, exact as Fractions."""
    pts = [[Fraction(x) for x in lb], [Fraction(x) for x in ub]]
    for _ in range(n_random):
        p = []
        for lo, hi in zip(lb, ub):
            x = lo + (hi - lo) * rng.random()
            p.append(Fraction(min(max(x, lo), hi)))
        pts.append(p)
    return pts


def _contains(lo, hi, v):
    lo_ok = lo == -math.inf or (not math.isnan(lo) and Fraction(lo) <= v.lo)
    hi_ok = hi == math.inf or (not math.isnan(hi) and v.hi <= Fraction(hi))
    return lo_ok and hi_ok


@pytest.mark.parametrize("key", KEYS)
def test_device_intervals_enclose_point_values(key):
    prog = sqd.Program.from_lemma(key)
    text = prog.text()
    pp = point_eval.Program(text)
    rng = random.Random(20261203)
    with sqd.Core(prog) as core:
        lb, ub = _boxes(core, text, rng)
        b64, c64 = core.eval_program(lb, ub, 64)
        b32, c32 = core.eval_program(lb, ub, 32)
    assert b64.shape[1] == len(pp.bounds) and c64.shape[1] == len(pp.constraints)
    roots = [b["expr"] for b in pp.bounds] + pp.constraints
    nb = len(pp.bounds)
    checked = skipped = 0
    for i in range(lb.shape[1]):
        for pt in _points(lb[:, i], ub[:, i], rng):
            vals = pp.evaluate(pt, roots)
            for prec, bv, cv in ((64, b64, c64), (32, b32, c32)):
                for k in range(nb):
                    v = vals[k]
                    lo, hi, und = bv[i, k]
                    if v is point_eval.UNKNOWN:
                        skipped += 1
                    elif v is point_eval.UNDEF:
                        assert und, (key, prec, "bound", k, "undefined at a point, not flagged", i)
                        checked += 1
                    else:
                        assert _contains(lo, hi, v), (key, prec, "bound", k, i, lo, hi, float(v.lo), float(v.hi))
                        checked += 1
                for j in range(len(pp.constraints)):
                    t = vals[nb + j]
                    definitely, possibly = cv[i, j]
                    if t is True:
                        assert possibly, (key, prec, "constraint", pp.constraints[j], "holds at a point", i)
                        checked += 1
                    elif t is False:
                        assert not definitely, (key, prec, "constraint", pp.constraints[j], "fails at a point", i)
                        checked += 1
                    else:
                        skipped += 1
    assert checked > 0 and skipped <= checked // 10, (checked, skipped)
