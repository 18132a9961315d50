"""Shared checks of the enclosure tests (test_gpu_enclosure.py on the device,
test_oracle_enclosure.py on the oracle): boxes the proofs visit, sampled points
of them, and the containment of the points' real values in the interval values.

The reference's tests evaluate its interval functions on random points and
intervals and require the interval results to contain the point results
(ivarp/test/test_util.hpp:113-235, test/cuda/cuda_test_fn.cuh:29-86).  Here the
functions are each statement's bound expressions and constraints, and the point
values come from tests/point_eval.py (near-exact arithmetic):

  bound k:       [lb, ub] contains the real value wherever it is defined, and is
                 flagged possibly undefined wherever it is not;
  constraint j:  a point where it holds has `possibly`, one where it fails has
                 not `definitely`.
"""
import math
import random
from fractions import Fraction

import numpy as np

import point_eval

KEYS = ["top1", "top2", "one", "two", "three", "four", "five", "six", "seven", "more1", "more2", "ec1", "ec2", "ec3"]
SEED = 20261203


def splits(text):
    out = []
    for line in text.splitlines():
        f = line.split()
        if f and f[0] == "split":
            out.append((int(f[1]), int(f[2]), int(f[3]), f[4] == "1"))
    return out


def boxes(initial_box, initial_queue, text, rng, n_queue=12):
    """SoA boxes (N, n): the initial box, initial-queue boxes and random cells of
    their static grids and dynamic subdivisions."""
    blb, bub = initial_box
    qlb, qub = initial_queue
    pick = rng.sample(range(qlb.shape[1]), min(n_queue, qlb.shape[1]))
    lbs = [blb[:, None], qlb[:, pick]]
    ubs = [bub[:, None], qub[:, pick]]
    for dynamic in (False, True):                  # one random cell of every static (dynamic) variable
        clb, cub = qlb[:, pick].copy(), qub[:, pick].copy()
        for arg, _, subs, dyn in splits(text):
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


def points(lb, ub, rng, n_random=3):
    """Points of a box: its two corners and random points, exact as Fractions."""
    pts = [[Fraction(x) for x in lb], [Fraction(x) for x in ub]]
    for _ in range(n_random):
        p = []
        for lo, hi in zip(lb, ub):
            x = lo + (hi - lo) * rng.random()
            p.append(Fraction(min(max(x, lo), hi)))
        pts.append(p)
    return pts


def contains(lo, hi, v):
    lo_ok = lo == -math.inf or (not math.isnan(lo) and Fraction(lo) <= v.lo)
    hi_ok = hi == math.inf or (not math.isnan(hi) and v.hi <= Fraction(hi))
    return lo_ok and hi_ok


def check(key, pp, lb, ub, evaluations, rng):
    """Asserts the enclosure of the point values of every box in `evaluations`
    ({precision: (bound values, constraint values)}); returns (checked, skipped)."""
    roots = [b["expr"] for b in pp.bounds] + pp.constraints
    nb = len(pp.bounds)
    checked = skipped = 0
    for i in range(lb.shape[1]):
        for pt in points(lb[:, i], ub[:, i], rng):
            vals = pp.evaluate(pt, roots)
            for prec, (bv, cv) in evaluations.items():
                for k in range(nb):
                    v = vals[k]
                    lo, hi, und = bv[i, k]
                    if v is point_eval.UNKNOWN:
                        skipped += 1
                    elif v is point_eval.UNDEF:
                        assert und, (key, prec, "bound", k, "undefined at a point, not flagged", i)
                        checked += 1
                    else:
                        assert contains(lo, hi, v), (key, prec, "bound", k, i, lo, hi, float(v.lo), float(v.hi))
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
    return checked, skipped


def oracle_boxes(oracle, text):
    """The boxes of `boxes` built from the oracle's initial box and queue (bit
    for bit the device's), with the seeded generator the tests use."""
    rng = random.Random(SEED)
    qlb, qub, _ = oracle.initial_queue()
    return boxes(oracle.initial_box(), (qlb, qub), text, rng), rng
