"""The binary32 acos of the FP32 prefilter (oracle restatement; the device's is
compared with it bit for bit in test_gpu_fp32.py): an enclosure made of the
arcsine series on binary64 intervals, converted outward to floats.  Pinned by
MPFR: it must contain acos rounded correctly down and up in binary64, and it
must almost always be the correctly rounded pair of floats (a wrong or missing
series coefficient, a wrong branch or a wrong sign breaks one of the two)."""
import math

import numpy as np
import pytest

import pyoracle


def _f32_down(v):
    f = np.float32(v)
    return np.float64(f) if np.float64(f) <= v else np.float64(np.nextafter(f, np.float32(-np.inf)))


def _f32_up(v):
    f = np.float32(v)
    return np.float64(f) if np.float64(f) >= v else np.float64(np.nextafter(f, np.float32(np.inf)))


def _inputs():
    rng = np.random.default_rng(20261203)
    x = np.concatenate([
        rng.uniform(-1, 1, 60000),
        1 - np.exp2(-rng.uniform(1, 24, 15000)),                  # near 1 (acos near 0)
        -1 + np.exp2(-rng.uniform(1, 24, 15000)),                 # near -1 (acos near pi)
        np.exp2(-rng.uniform(1, 126, 8000)) * rng.choice([-1, 1], 8000),   # near 0
        0.5 + rng.uniform(-1e-3, 1e-3, 4000),                     # the branch point
        -0.5 + rng.uniform(-1e-3, 1e-3, 4000),
        [0.5, -0.5, 1.0, -1.0, 0.0, -0.0, 2.0 ** -149, -(2.0 ** -149)],
    ])
    x = x.astype(np.float32).astype(np.float64)
    x[np.abs(x) > 1] = np.sign(x[np.abs(x) > 1])
    return x


def test_contains_correctly_rounded_acos_and_is_tight():
    x = _inputs()
    lo, hi = pyoracle.rounded_op("acos32", x, x)
    elo, ehi = pyoracle.rounded_op("acos", x, x)                   # MPFR, binary64, RD / RU
    assert np.all(lo <= elo) and np.all(hi >= ehi)
    clo = np.array([_f32_down(v) for v in elo])
    chi = np.array([_f32_up(v) for v in ehi])
    assert np.all(lo <= clo) and np.all(hi >= chi)                  # the floats around acos
    tight = np.mean((lo == clo) & (hi == chi))
    assert tight > 0.999, tight
    assert np.all(np.float64(np.float32(lo)) == lo) and np.all(np.float64(np.float32(hi)) == hi)


@pytest.mark.parametrize("t,value", [(1.0, 0.0), (0.0, math.pi / 2), (0.5, math.pi / 3), (-0.5, 2 * math.pi / 3),
                                     (-1.0, math.pi), (math.sqrt(0.5), math.pi / 4)])
def test_known_values(t, value):
    t = float(np.float32(t))
    lo, hi = pyoracle.rounded_op("acos32", np.array([t]), np.array([t]))
    if t == 1.0:
        assert lo[0] == 0.0 and hi[0] == 0.0
        return
    exact = value if t in (0.5, -0.5, -1.0, 0.0) else math.acos(t)
    assert lo[0] < exact < hi[0]                                    # irrational: strictly inside
    assert hi[0] - lo[0] <= 2 * float(np.spacing(np.float32(exact)))


def test_outside_the_domain_is_nan():
    lo, hi = pyoracle.rounded_op("acos32", np.array([1.5, -2.0, np.inf, np.nan]), np.zeros(4))
    assert np.all(np.isnan(lo)) and np.all(np.isnan(hi))


def test_acos_of_one_is_positive_zero():
    lo, hi = pyoracle.rounded_op("acos32", np.array([1.0]), np.array([1.0]))
    assert np.signbit(lo[0]) == False and np.signbit(hi[0]) == False
