"""Device interval operations (sqd_interval_op, the same device functions the
prover kernels use with run-time sign decisions) against the reference's
known-answer interval tests and bit for bit against the CPU oracle on random
intervals, including zero-containing, infinite and undefined operands."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import pyoracle
import sqd

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "interval_kat.json").read_text())
INF = float("inf")


def dec(v):
    return INF if v == "inf" else (-INF if v == "-inf" else float.fromhex(v))


def iv(t):
    return [dec(t[0]), dec(t[1]), float(t[2])]


def test_device_interval_known_answers():
    by_op = {}
    for c in GOLDEN["cases"]:
        by_op.setdefault(c["op"], []).append(c)
    for op, cases in by_op.items():
        got = sqd.interval_op(op, [iv(c["a"]) for c in cases], [iv(c["b"]) for c in cases])
        for c, g in zip(cases, got):
            e = iv(c["expect"])
            assert g[0] == e[0] and g[1] == e[1] and bool(g[2]) == bool(e[2]), (c["src"], g, e)


def _random_intervals(rng, n):
    vals = np.concatenate([rng.uniform(-4, 4, 8) * np.exp2(rng.integers(-30, 30, 8)),
                           [0.0, -0.0, 1.0, -1.0, INF, -INF, 5e-324, -5e-324, 1.7976931348623157e308]])
    a = rng.choice(vals, n)
    b = rng.choice(vals, n)
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    mixed = rng.uniform(-100, 100, n)
    w = rng.uniform(0, 50, n)
    use = rng.random(n) < 0.5
    lo = np.where(use, mixed, lo)
    hi = np.where(use, mixed + w, hi)
    und = (rng.random(n) < 0.1).astype(float)
    return np.stack([lo, hi, und], axis=1)


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "sqrt", "square", "neg", "max", "min"])
def test_device_interval_ops_match_oracle(op):
    rng = np.random.default_rng(77 + len(op))
    A = _random_intervals(rng, 200000)
    B = _random_intervals(rng, 200000)
    g = sqd.interval_op(op, A, B)
    o = pyoracle.interval_op(op, A, B)
    nan_g, nan_o = np.isnan(g), np.isnan(o)
    assert np.array_equal(nan_g, nan_o)
    assert np.array_equal(_bits(g)[~nan_o], _bits(o)[~nan_o])


def test_device_splitter_matches_oracle():
    rng = np.random.default_rng(3)
    lo = rng.uniform(-2, 2, 5000)
    hi = lo + rng.uniform(0, 3, 5000) * np.exp2(rng.integers(-40, 4, 5000))
    n = rng.integers(1, 300, 5000).astype(float)
    i = np.floor(rng.random(5000) * (n + 1))
    A = np.stack([lo, hi, np.zeros(5000)], axis=1)
    B = np.stack([i, n, np.zeros(5000)], axis=1)
    g = sqd.interval_op("split", A, B)[:, 0]
    o = pyoracle.interval_op("split", A, B)[:, 0]
    assert np.array_equal(_bits(g), _bits(o))
