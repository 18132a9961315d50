"""GPU prover core (through the C ABI) against the CPU oracle: bit-exact
arithmetic, initial box and queue, non-final/final passes, dynamic split and
whole proofs."""
import numpy as np
import pytest

import pyoracle
import sqd
from conftest import PUBLISHED

pytestmark = pytest.mark.gpu

GPU_LEMMAS = ["top1", "top2", "one", "two", "three", "four", "five", "six", "seven", "more1", "more2"]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _edge_values():
    v = [0.0, -0.0, 1.0, -1.0, 0.5, 2.0, 3.0, 5.0, 1e-300, -1e-300, 1e300, -1e300, 5e-324, -5e-324,
         2.2250738585072014e-308, 3e-310, -7.5e-321, 1e10, -3e15, 1.7976931348623157e308, np.inf, -np.inf,
         1.2649110640673517, 0.295, 1.6, 0.1, 1 / 3]
    return np.array(v)


def _inputs(seed, n):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-4, 4, n) * np.exp2(rng.integers(-20, 20, n))
    b = rng.uniform(-4, 4, n) * np.exp2(rng.integers(-20, 20, n))
    e = _edge_values()
    ea, eb = np.meshgrid(e, e)
    return np.concatenate([a, ea.ravel()]), np.concatenate([b, eb.ravel()])


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "sqrt"])
def test_rounded_ops_bit_exact(op):
    a, b = _inputs(20261202 + len(op), 1 << 20)
    if op == "sqrt":
        a = np.abs(a)
    glo, ghi = sqd.rounded_op(op, a, b)
    clo, chi = pyoracle.rounded_op(op, a, b)
    nan = np.isnan(clo)
    assert np.array_equal(np.isnan(glo), nan) and np.array_equal(np.isnan(ghi), np.isnan(chi))
    assert np.array_equal(_bits(glo)[~nan], _bits(clo)[~nan])
    assert np.array_equal(_bits(ghi)[~np.isnan(chi)], _bits(chi)[~np.isnan(chi)])


def _acos_inputs(seed, n):
    rng = np.random.default_rng(seed)
    x = np.concatenate([
        rng.uniform(-1, 1, n),
        1 - np.ldexp(rng.integers(1, 1 << 20, n).astype(float), -53),            # just below 1
        -1 + np.ldexp(rng.integers(1, 1 << 20, n // 4).astype(float), -53),      # just above -1
        # 1 - k^2/2 * 2^(2j-53): sqrt(2 (1 - x)) is a short dyadic number (hardest to round)
        1 - np.ldexp(rng.integers(1, 64, 4096).astype(float) ** 2 / 2, -53 + 2 * rng.integers(0, 20, 4096)),
        np.ldexp(rng.uniform(0.5, 1, n), -rng.integers(0, 200, n)),              # tiny magnitudes
        [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 5e-324, -5e-324, 2.2250738585072014e-308]])
    return x[(x >= -1) & (x <= 1)]


def test_acos_bit_exact_with_mpfr():
    # the reference rounds acos with MPFR (ivarp/src/interval_acos.cpp:29-66); the oracle calls MPFR
    x = _acos_inputs(5, 1 << 18)
    before = sqd.acos_fallbacks()
    glo, ghi = sqd.rounded_op("acos", x, x)
    olo, ohi = pyoracle.rounded_op("acos", x, x)
    assert np.array_equal(_bits(glo), _bits(olo)) and np.array_equal(_bits(ghi), _bits(ohi))
    assert sqd.acos_fallbacks() == before        # every direction decided at 2^-113
    out = sqd.rounded_op("acos", np.array([1.5, -2.0, np.inf, np.nan]), np.zeros(4))
    assert np.all(np.isnan(out[0])) and np.all(np.isnan(out[1]))


@pytest.mark.parametrize("op", ["acos", "acos_series"])
def test_acos_paths_bit_exact_with_mpfr(op):
    # both device methods: the cos walk (with the series as its fallback) and the series alone
    x = _acos_inputs(17, 1 << 16)
    before = sqd.acos_fallbacks()
    glo, ghi = sqd.rounded_op(op, x, x)
    olo, ohi = pyoracle.rounded_op(op, x, x)
    assert np.array_equal(_bits(glo), _bits(olo)) and np.array_equal(_bits(ghi), _bits(ohi))
    assert sqd.acos_fallbacks() == before


def test_acos_interval_domain_clamp():
    # interval_asin.hpp:13-27: with unbounded compile-time bounds the argument is clamped to
    # [-1, 1] and marked possibly undefined
    A = np.array([[-2.0, 0.5, 0.0], [-0.25, 3.0, 0.0], [0.1, 0.2, 0.0], [0.3, 0.4, 1.0]])
    g = sqd.interval_op("acos", A)
    o = pyoracle.interval_op("acos", A)
    assert np.array_equal(_bits(g), _bits(o))
    assert list(g[:, 2]) == [1.0, 1.0, 0.0, 1.0]


@pytest.fixture(scope="module")
def cores():
    made = {}

    def get(key, generations=8):
        k = (key, generations)
        if k not in made:
            p = sqd.Program.from_lemma(key)
            made[k] = (p, sqd.Core(p, settings=sqd.default_settings(generation_count=generations)))
        return made[k][1]
    yield get
    for p, c in made.values():
        c.close()


@pytest.mark.parametrize("key", GPU_LEMMAS + ["ec1", "ec2"])
def test_initial_box(cores, program_text, key):
    glb, gub = cores(key).initial_box()
    olb, oub = pyoracle.Oracle(program_text(key)).initial_box()
    assert np.array_equal(_bits(glb), _bits(olb)) and np.array_equal(_bits(gub), _bits(oub))


def _sorted_boxes(lb, ub):
    m = np.vstack([_bits(lb), _bits(ub)]).T
    return m[np.lexsort(m.T[::-1])]


@pytest.mark.parametrize("key", ["one", "two", "three", "five", "more2", "ec2"])
def test_initial_queue(cores, program_text, key):
    glb, gub, gc = cores(key).initial_queue()
    olb, oub, oc = pyoracle.Oracle(program_text(key)).initial_queue()
    assert gc["num_cuboids"] == oc["num_cuboids"] and gc["num_leaf_cuboids"] == oc["num_leaf_cuboids"]
    assert np.array_equal(_sorted_boxes(glb, gub), _sorted_boxes(olb, oub))


@pytest.mark.parametrize("key,limit", [("one", 128), ("two", 512), ("three", 256), ("four", 128), ("five", 256),
                                       ("six", 128), ("seven", 64), ("more1", 256), ("more2", 64), ("ec2", 16),
                                       ("ec3", 4)])
def test_nonfinal_hulls(cores, program_text, key, limit):
    o = pyoracle.Oracle(program_text(key))
    qlb, qub, _ = o.initial_queue()
    idx = np.linspace(0, qlb.shape[1] - 1, min(limit, qlb.shape[1])).astype(int)
    lb, ub = qlb[:, idx], qub[:, idx]
    ghl, ghu, ghas, gc = cores(key).nonfinal(lb, ub)
    ohl, ohu, ohas, oc = o.nonfinal(lb, ub)
    assert gc["num_cuboids"] == oc["num_cuboids"] and gc["num_leaf_cuboids"] == oc["num_leaf_cuboids"]
    assert np.array_equal(ghas, ohas)
    keep = ohas.astype(bool)
    assert np.array_equal(_bits(ghl[:, keep]), _bits(ohl[:, keep]))
    assert np.array_equal(_bits(ghu[:, keep]), _bits(ohu[:, keep]))


@pytest.mark.parametrize("key,limit", [("one", 128), ("two", 256), ("four", 64), ("six", 64)])
def test_split_children(cores, program_text, key, limit):
    o = pyoracle.Oracle(program_text(key))
    qlb, qub, _ = o.initial_queue()
    idx = np.linspace(0, qlb.shape[1] - 1, min(limit, qlb.shape[1])).astype(int)
    hl, hu, has, _ = o.nonfinal(qlb[:, idx], qub[:, idx])
    hl, hu = hl[:, has.astype(bool)], hu[:, has.astype(bool)]
    glb, gub, gc = cores(key).split(hl, hu)
    olb, oub, oc = o.split(hl, hu)
    assert gc["num_cuboids"] == oc["num_cuboids"] and gc["num_leaf_cuboids"] == oc["num_leaf_cuboids"]
    assert np.array_equal(_sorted_boxes(glb, gub), _sorted_boxes(olb, oub))


def test_final_pass_direct(cores):
    # ivarp/test/expected_criticals2.cpp:99-148
    core = cores("ec2")
    for k in (4, 16, 64):
        pts = np.array([i / k for i in range(k + 1)])
        w = np.repeat(np.arange(k), k)
        x = np.tile(np.arange(k), k)
        n = k * k
        lb = np.vstack([pts[w], pts[x], np.zeros(n), np.zeros(n)])
        ub = np.vstack([pts[w + 1], pts[x + 1], np.ones(n), np.ones(n)])
        _, _, has, c = core.nonfinal(lb, ub)
        assert has.all() and c["num_leaf_cuboids"] == n * 64 and c["num_cuboids"] == n * 8 + n * 64
        cl, cu, cf = core.final(lb, ub, capacity=n * 64)
        assert cf["num_critical_cuboids"] == n * 64 and cl.shape[1] == n * 64


@pytest.mark.parametrize("key", GPU_LEMMAS)
def test_prove_published_counts(cores, key):
    core = cores(key)
    r = core.prove()
    c = r["counts"]
    assert (c["num_cuboids"], c["num_leaf_cuboids"]) == PUBLISHED[key]
    assert c["num_critical_cuboids"] == 0 and r["n_critical"] == 0
    assert core.acos_fallbacks() == 0


def test_prove_expected_criticals_2d(cores, program_text):
    r = cores("ec1", generations=2).prove(critical_capacity=1024)
    o_counts, o_crit = pyoracle.Oracle(program_text("ec1")).run(generations=2, threads=4, max_criticals=1024)
    assert r["counts"] == o_counts
    g = np.array(sorted(tuple(x) for x in np.vstack([r["crit_lb"], r["crit_ub"]]).T[:, [0, 2, 1, 3]]))
    o = np.array(sorted(tuple(b.reshape(-1)) for b in o_crit))
    assert np.array_equal(g.view(np.uint64), o.view(np.uint64))


@pytest.mark.parametrize("gens", [0, 1, 2])
def test_prove_expected_criticals_4d(cores, program_text, gens):
    r = cores("ec2", generations=gens).prove(critical_capacity=1 << 19)
    o_counts, o_crit = pyoracle.Oracle(program_text("ec2")).run(generations=gens, threads=8, max_criticals=1 << 19)
    assert r["counts"] == o_counts
    g = np.vstack([r["crit_lb"], r["crit_ub"]]).T[:, [0, 4, 1, 5, 2, 6, 3, 7]]
    o = o_crit.reshape(len(o_crit), -1)
    gs = g[np.lexsort(g.T[::-1])]
    os_ = o[np.lexsort(o.T[::-1])]
    assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64))


@pytest.mark.parametrize("gens", [0, 1, 2])
def test_prove_expected_criticals_hoisted(cores, program_text, gens):
    # ec3: the last static level splits z 128 ways and is decided per parent (its
    # terms free of z hoisted), the children whose z range raises y's lower bound
    # apart from their siblings; counts and every critical box are the oracle's,
    # on the first proof of the core and on a replay of its round graphs
    o_counts, o_crit = pyoracle.Oracle(program_text("ec3")).run(generations=gens, threads=8, max_criticals=1 << 16)
    o = o_crit.reshape(len(o_crit), -1)
    os_ = o[np.lexsort(o.T[::-1])]
    core = cores("ec3", generations=gens)
    for _ in range(2):
        r = core.prove(critical_capacity=1 << 16)
        assert r["counts"] == o_counts and r["n_critical"] == len(o_crit) > 0
        g = np.vstack([r["crit_lb"], r["crit_ub"]]).T[:, [0, 3, 1, 4, 2, 5]]
        gs = g[np.lexsort(g.T[::-1])]
        assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64))
