"""The binary64 group test (sqd_settings.group_levels) on the device, through the
C ABI: it only changes which children are evaluated one by one, so non-final
hulls, critical boxes and counts are bit for bit those of the plain binary64
pass (the oracle's), for several group sizes."""
import numpy as np
import pytest

import pyoracle
import sqd
from conftest import PUBLISHED

pytestmark = pytest.mark.gpu

ALL = 0xFFFFFFFF


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("key", ["top1", "top2", "one", "two", "three", "four", "five", "six", "seven", "more2"])
@pytest.mark.parametrize("group", [2, 16, 64])
def test_grouped_prove_published(key, group):
    p = sqd.Program.from_lemma(key)
    with sqd.Core(p, settings=sqd.default_settings(group_levels=ALL, prefilter_group=group)) as core:
        r = core.prove()
    c = r["counts"]
    assert (c["num_cuboids"], c["num_leaf_cuboids"]) == PUBLISHED[key]
    assert c["num_critical_cuboids"] == 0 and r["n_critical"] == 0


@pytest.mark.parametrize("key,limit", [("one", 256), ("two", 512), ("three", 256), ("four", 64), ("more2", 32),
                                       ("ec2", 16)])
@pytest.mark.parametrize("group", [4, 32])
def test_grouped_nonfinal_hulls_bit_exact(program_text, key, limit, group):
    o = pyoracle.Oracle(program_text(key))
    qlb, qub, _ = o.initial_queue()
    idx = np.linspace(0, qlb.shape[1] - 1, min(limit, qlb.shape[1])).astype(int)
    lb, ub = qlb[:, idx], qub[:, idx]
    ohl, ohu, ohas, oc = o.nonfinal(lb, ub)
    p = sqd.Program.from_lemma(key)
    with sqd.Core(p, settings=sqd.default_settings(group_levels=ALL, prefilter_group=group)) as core:
        ghl, ghu, ghas, gc = core.nonfinal(lb, ub)
    assert gc["num_cuboids"] == oc["num_cuboids"] and gc["num_leaf_cuboids"] == oc["num_leaf_cuboids"]
    assert np.array_equal(ghas, ohas)
    keep = ohas.astype(bool)
    assert np.array_equal(_bits(ghl[:, keep]), _bits(ohl[:, keep]))
    assert np.array_equal(_bits(ghu[:, keep]), _bits(ohu[:, keep]))


@pytest.mark.parametrize("key", ["ec2", "ec3"])
@pytest.mark.parametrize("gens", [0, 1, 2])
def test_grouped_criticals_bit_exact(program_text, key, gens):
    p = sqd.Program.from_lemma(key)
    s = sqd.default_settings(generation_count=gens, group_levels=ALL, prefilter_group=8)
    with sqd.Core(p, settings=s) as core:
        r = core.prove(critical_capacity=1 << 19)
    o = pyoracle.Oracle(program_text(key))
    oc, ocrit = o.run(generations=gens, threads=8, max_criticals=1 << 19)
    assert r["counts"] == oc
    nv = r["crit_lb"].shape[0]
    g = np.vstack([r["crit_lb"], r["crit_ub"]]).T[:, [k for a in range(nv) for k in (a, nv + a)]]
    oo = ocrit.reshape(len(ocrit), -1)
    assert np.array_equal(g[np.lexsort(g.T[::-1])].view(np.uint64), oo[np.lexsort(oo.T[::-1])].view(np.uint64))


def test_grouped_with_prefilter_and_batches():
    # the prefilter on level 1, the binary64 group test on level 0, in batches
    p = sqd.Program.from_lemma("one")
    s = sqd.default_settings(fp32_prefilter=2, group_levels=1, prefilter_group=16, jobs=4)
    with sqd.Core(p, settings=s) as core:
        n, _ = core.shape()
        got = core.prove_batch([(0, n), (0, n // 2), (n // 2, n)])
    assert (got[0]["counts"]["num_cuboids"], got[0]["counts"]["num_leaf_cuboids"]) == PUBLISHED["one"]
    assert got[1]["counts"]["num_cuboids"] + got[2]["counts"]["num_cuboids"] == PUBLISHED["one"][0]


def test_grouped_small_buffers_exact():
    # small level buffers and the largest round size: the core shrinks its rounds
    # so that a grouped level's children always fit the undecided list
    p = sqd.Program.from_lemma("two")
    s = sqd.default_settings(group_levels=ALL, prefilter_group=64, level_bytes=4 << 20, dequeue_buffer_size=65536)
    with sqd.Core(p, settings=s) as core:
        r = core.prove()
    assert (r["counts"]["num_cuboids"], r["counts"]["num_leaf_cuboids"]) == PUBLISHED["two"]


def test_prefilter_small_buffers_exact():
    # with small buffers and a large round size the core bounds its rounds so
    # that the undecided list of every prefiltered level holds all of its
    # children: the proof needs no retry and keeps the published counts
    p = sqd.Program.from_lemma("two")
    s = sqd.default_settings(fp32_prefilter=sqd.FP32_ALL_LEVELS, level_bytes=4 << 20, dequeue_buffer_size=65536)
    with sqd.Core(p, settings=s) as core:
        r = core.prove()
        assert (r["counts"]["num_cuboids"], r["counts"]["num_leaf_cuboids"]) == PUBLISHED["two"]
        r = core.prove_range(5, 6)
        assert r["counts"]["num_cuboids"] > 0
