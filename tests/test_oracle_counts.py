"""The CPU oracle against the reference's published totals and the known-answer
prover tests of ivarp (expected_criticals.cpp, expected_criticals2.cpp)."""
import numpy as np
import pytest
from fractions import Fraction

import pyoracle
from conftest import PUBLISHED


@pytest.mark.parametrize("key", ["top1", "top2", "one", "two", "three", "five", "six", "seven"])
def test_oracle_published_counts(program_text, key):
    counts, _ = pyoracle.Oracle(program_text(key)).run(generations=8, threads=8)
    assert (counts["num_cuboids"], counts["num_leaf_cuboids"]) == PUBLISHED[key]
    assert counts["num_critical_cuboids"] == 0


def test_oracle_published_counts_four(program_text):
    counts, _ = pyoracle.Oracle(program_text("four")).run(generations=8, threads=8)
    assert (counts["num_cuboids"], counts["num_leaf_cuboids"]) == PUBLISHED["four"]


def _expected_ec1_boxes():
    # closed form of ivarp/test/expected_criticals.cpp:32-50 (IDouble)
    boxes = []
    for xi in range(32):
        xl = float(Fraction(xi, 16))
        xu = float(Fraction(xi + 1, 16))  # multiples of 1/16 are exact
        ymax = min(xu, 3 - xl)
        for yj in range(4):
            miny = (ymax / 4) * yj   # exact: division by 4, small integer multiple
            maxy = (ymax / 4) * (yj + 1)
            lo = xl
            if miny > 0 and lo < miny:
                lo = miny
            boxes.append((lo, xu, miny, maxy))
    return sorted(boxes)


def test_oracle_expected_criticals_2d(program_text):
    o = pyoracle.Oracle(program_text("ec1"))
    counts, crit = o.run(generations=2, threads=4, max_criticals=1000)
    assert counts["num_critical_cuboids"] == 128
    assert counts["num_cuboids"] == (8 + 8 * 4) + (16 + 16 * 4) + (32 + 32 * 4) + 32 * 4
    assert counts["num_leaf_cuboids"] == 8 * 4 + 16 * 4 + 2 * 32 * 4
    got = sorted(tuple(float(v) for v in (b[0, 0], b[0, 1], b[1, 0], b[1, 1])) for b in crit)
    assert got == _expected_ec1_boxes()


@pytest.mark.parametrize("gens,cuboids,leaves,crits", [
    (0, 1172 + 1152, 2 * 1024, 1024),
    (1, 4 + 16 + 8 * 16 + 64 * 16 + 4 * 16 + 4 * 4 * 16 + 256 * 8 + 256 * 64 + 256 * 8 + 256 * 64, 1024 + 2 * 16384, 16384),
    (2, 614868, 1024 + 16384 + 262144 * 2, 262144),
])
def test_oracle_expected_criticals_4d(program_text, gens, cuboids, leaves, crits):
    o = pyoracle.Oracle(program_text("ec2"))
    counts, crit = o.run(generations=gens, threads=8, max_criticals=crits)
    assert counts["num_critical_cuboids"] == crits
    assert counts["num_cuboids"] == cuboids
    assert counts["num_leaf_cuboids"] == leaves
    if gens == 2:
        # criticals are the full product grid 64 x 64 x 8 x 8 of the unit cube
        # (ivarp/test/expected_criticals2.cpp:42-62); splitter points of [0,1] into
        # 2^k parts are exact dyadic numbers
        def pts(n):
            return [i / n for i in range(n + 1)]
        grid = np.array(sorted((w0, w1, x0, x1, y0, y1, z0, z1)
                               for w0, w1 in zip(pts(64)[:-1], pts(64)[1:])
                               for x0, x1 in zip(pts(64)[:-1], pts(64)[1:])
                               for y0, y1 in zip(pts(8)[:-1], pts(8)[1:])
                               for z0, z1 in zip(pts(8)[:-1], pts(8)[1:])))
        got = np.array(sorted(tuple(crit[i].reshape(-1)) for i in range(len(crit))))
        assert np.array_equal(got, grid)


def test_oracle_core_counts_direct(program_text):
    # ivarp/test/expected_criticals2.cpp:99-148: the core driven directly
    o = pyoracle.Oracle(program_text("ec2"))
    for k in (4, 16, 64):
        pts = np.array([i / k for i in range(k + 1)])
        w = np.repeat(np.arange(k), k)
        x = np.tile(np.arange(k), k)
        n = k * k
        lb = np.vstack([pts[w], pts[x], np.zeros(n), np.zeros(n)])
        ub = np.vstack([pts[w + 1], pts[x + 1], np.ones(n), np.ones(n)])
        _, _, has, c = o.nonfinal(lb, ub)
        assert has.all()
        assert c["num_leaf_cuboids"] == n * 8 * 8
        assert c["num_cuboids"] == n * 8 + n * 8 * 8
        cl, cu, cf = o.final(lb, ub, cap=n * 64)
        assert cf["num_critical_cuboids"] == n * 64 and cf["num_leaf_cuboids"] == n * 64
        assert cf["num_cuboids"] == n * 8 + n * 8 * 8


@pytest.mark.parametrize("key", ["one", "three"])
def test_oracle_fp32_prefilter_keeps_published_counts(program_text, key):
    # With the FP32 prefilter the oracle's counts stay the published ones (every
    # child binary32 prunes is one binary64 prunes too, for these statements) and
    # binary32 decides the large majority of static-level children.
    o = pyoracle.Oracle(program_text(key), fp32=True)
    counts, _ = o.run(threads=8)
    assert (counts["num_cuboids"], counts["num_leaf_cuboids"]) == PUBLISHED[key]
    assert counts["num_critical_cuboids"] == 0
    checked, decided = o.fp32_stats()
    assert checked > 0 and 0.95 < decided / checked <= 1.0


def test_oracle_init_range_partition_adds_up(program_text):
    # a job's range of the first variable's initial subdivisions is an exact
    # union of subtrees: a partition of [0, 128) adds up to the published totals
    tot = [0, 0]
    for a, b in [(0, 40), (40, 64), (64, 100), (100, 128)]:
        o = pyoracle.Oracle(program_text("one"))
        o.set_init_range(a, b)
        c, _ = o.run(threads=8)
        tot[0] += c["num_cuboids"]
        tot[1] += c["num_leaf_cuboids"]
    assert tuple(tot) == PUBLISHED["one"]
