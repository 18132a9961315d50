"""The FP32 prefilter on the device against the oracle's restatement of it
(oracle_set_fp32_prefilter), through the C ABI: binary32 directed arithmetic bit
for bit, non-final hulls, critical boxes and whole-proof counts, plus the
verdicts of every statement (zero criticals, the published counts)."""
import numpy as np
import pytest

import pyoracle
import sqd
from conftest import PUBLISHED

pytestmark = pytest.mark.gpu

LEMMAS = ["top1", "top2", "one", "two", "three", "four", "five", "six", "seven", "more1", "more2"]


def _floats(seed, n):
    rng = np.random.default_rng(seed)
    m = rng.uniform(0.5, 1.0, n) * np.exp2(rng.integers(-60, 60, n).astype(float))
    v = (m * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    edge = np.array([0.0, -0.0, 1.0, -1.0, 2.0, 0.5, 3.0, 1e-45, -1e-45, 1e-38, 1.17549435e-38, 3.4028235e38,
                     -3.4028235e38, 1e-30, 4.0, 9.0, 16.0, np.inf, -np.inf], dtype=np.float32)
    return np.concatenate([v, edge]).astype(np.float64)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("op", ["add32", "sub32", "mul32", "div32", "sqrt32", "to32", "acos32"])
def test_binary32_ops_bit_exact(op):
    a = _floats(5, 200000)
    b = _floats(6, 200000)
    if op == "sqrt32":
        a = np.abs(a)
    if op == "acos32":                                # the prefilter's enclosure (tests/test_acos_enclosure.py)
        rng = np.random.default_rng(7)
        a = np.concatenate([rng.uniform(-1, 1, 100000), 1 - np.exp2(-rng.uniform(1, 24, 20000)),
                            -1 + np.exp2(-rng.uniform(1, 24, 20000)),
                            np.exp2(-rng.uniform(1, 126, 10000)) * rng.choice([-1, 1], 10000),
                            0.5 + rng.uniform(-1e-3, 1e-3, 5000), -0.5 + rng.uniform(-1e-3, 1e-3, 5000),
                            [1.0, -1.0, 0.0, -0.0, 0.5, -0.5, 1.5, -2.0]])
        a = a.astype(np.float32).astype(np.float64)
        b = a
    if op == "to32":
        a = np.random.default_rng(8).standard_normal(200000) * np.exp2(np.random.default_rng(9).integers(-140, 130, 200000))
        b = a
    glo, ghi = sqd.rounded_op(op, a, b)
    olo, ohi = pyoracle.rounded_op(op, a, b)
    for g, o in ((glo, olo), (ghi, ohi)):          # NaN results (0/0, inf-inf) compare as NaN
        nan = np.isnan(o)
        assert np.array_equal(np.isnan(g), nan)
        assert np.array_equal(_bits(g[~nan]), _bits(o[~nan]))


@pytest.fixture(scope="module")
def fp32_cores():
    made = {}

    def get(key, generations=8):
        k = (key, generations)
        if k not in made:
            p = sqd.Program.from_lemma(key)
            s = sqd.default_settings(generation_count=generations, fp32_prefilter=sqd.FP32_ALL_LEVELS)
            made[k] = (p, sqd.Core(p, settings=s))
        return made[k][1]
    yield get
    for p, c in made.values():
        c.close()


@pytest.mark.parametrize("key,limit", [("one", 128), ("two", 512), ("three", 256), ("four", 64), ("more1", 128),
                                       ("more2", 32), ("ec2", 16)])
def test_fp32_nonfinal_hulls(fp32_cores, program_text, key, limit):
    o = pyoracle.Oracle(program_text(key), fp32=True)
    qlb, qub, _ = o.initial_queue()
    idx = np.linspace(0, qlb.shape[1] - 1, min(limit, qlb.shape[1])).astype(int)
    lb, ub = qlb[:, idx], qub[:, idx]
    core = fp32_cores(key)
    ghl, ghu, ghas, gc = core.nonfinal(lb, ub)
    ohl, ohu, ohas, oc = o.nonfinal(lb, ub)
    assert gc["num_cuboids"] == oc["num_cuboids"] and gc["num_leaf_cuboids"] == oc["num_leaf_cuboids"]
    assert core.fp32_stats() == o.fp32_stats()
    assert np.array_equal(ghas, ohas)
    keep = ohas.astype(bool)
    assert np.array_equal(_bits(ghl[:, keep]), _bits(ohl[:, keep]))
    assert np.array_equal(_bits(ghu[:, keep]), _bits(ohu[:, keep]))


@pytest.mark.parametrize("key", ["one", "two", "three"])
def test_fp32_prove_matches_oracle(fp32_cores, program_text, key):
    core = fp32_cores(key)
    r = core.prove()
    o = pyoracle.Oracle(program_text(key), fp32=True)
    oc, _ = o.run(threads=8)
    assert r["counts"] == oc
    assert core.fp32_stats() == o.fp32_stats()


@pytest.mark.parametrize("key", LEMMAS)
def test_fp32_prove_verdicts(fp32_cores, key):
    # every statement is still proven (zero criticals), with the published counts
    core = fp32_cores(key)
    r = core.prove()
    c = r["counts"]
    assert c["num_critical_cuboids"] == 0 and r["n_critical"] == 0
    assert (c["num_cuboids"], c["num_leaf_cuboids"]) == PUBLISHED[key]
    checked, decided = core.fp32_stats()
    assert decided <= checked


@pytest.mark.parametrize("key", ["ec2", "ec3"])
@pytest.mark.parametrize("gens", [0, 1, 2])
def test_fp32_expected_criticals(fp32_cores, program_text, key, gens):
    r = fp32_cores(key, generations=gens).prove(critical_capacity=1 << 19)
    o = pyoracle.Oracle(program_text(key), fp32=True)
    o_counts, o_crit = o.run(generations=gens, threads=8, max_criticals=1 << 19)
    assert r["counts"] == o_counts
    nv = r["crit_lb"].shape[0]
    g = np.vstack([r["crit_lb"], r["crit_ub"]]).T[:, [k for a in range(nv) for k in (a, nv + a)]]
    oo = o_crit.reshape(len(o_crit), -1)
    gs = g[np.lexsort(g.T[::-1])]
    os_ = oo[np.lexsort(oo.T[::-1])]
    assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64))


def test_width_stats_one(program_text):
    # binary32 enclosures are wider than binary64 ones on the prover's own boxes
    o = pyoracle.Oracle(program_text("one"))
    qlb, qub, _ = o.initial_queue()
    rng = np.random.default_rng(3)
    n = 4096
    pick = rng.integers(0, qlb.shape[1], n)
    lb, ub = qlb[:, pick].copy(), qub[:, pick].copy()
    for a in (1, 2):                                  # a random cell of the 256 x 256 static grid
        cell = rng.integers(0, 256, n)
        w = (ub[a] - lb[a]) / 256
        lb[a], ub[a] = lb[a] + cell * w, lb[a] + (cell + 1) * w
    p = sqd.Program.from_lemma("one")
    with sqd.Core(p) as core:
        st = core.width_stats(lb, ub)
    measured = [s for s in st if s["samples"] > 0]
    assert measured and all(s["mean_ratio"] >= 1.0 - 1e-9 for s in measured)
