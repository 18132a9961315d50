"""Batched proofs (sqd_prove_batch) on the device, through the C ABI: every job of
a batch gives exactly what the same range gives proved alone (counts, criticals
and FP32 prefilter counts), for full, partial and repeated batches, with graph
replays, and host-driven calls on a batch core still match the oracle."""
import numpy as np
import pytest

import pyoracle
import sqd
from conftest import PUBLISHED

pytestmark = pytest.mark.gpu


def _ranges(rng, n, k):
    out = []
    for _ in range(k):
        a = int(rng.integers(0, n))
        b = int(rng.integers(a + 1, n + 1))
        out.append((a, b))
    return out


@pytest.mark.parametrize("key", ["top1", "one", "two", "three", "four", "five", "six", "seven", "more2"])
@pytest.mark.parametrize("fp32", [0, sqd.FP32_ALL_LEVELS])
def test_batch_jobs_equal_single_ranges(key, fp32):
    rng = np.random.default_rng(len(key) * 7 + (fp32 & 1))
    p = sqd.Program.from_lemma(key)
    with sqd.Core(p, settings=sqd.default_settings(fp32_prefilter=fp32)) as single, \
            sqd.Core(p, settings=sqd.default_settings(fp32_prefilter=fp32, jobs=8)) as batch:
        n, _ = single.shape()
        if key == "more2":
            ranges = [(a, a + 1) for a in rng.integers(0, n, 8)]      # keep the run short
        else:
            ranges = _ranges(rng, n, 6) + [(0, n), (0, n)]              # two identical full proofs
        want = []
        for a, b in ranges:
            r = single.prove_range(a, b)
            want.append((r["counts"], r["n_critical"], single.fp32_stats()))
        got = batch.prove_batch(ranges)
        for (a, b), (wc, wn, wf), g in zip(ranges, want, got):
            assert g["counts"] == wc, (key, a, b)
            assert g["n_critical"] == wn == 0
            assert g["fp32_stats"] == wf, (key, a, b)
        for g, (a, b) in zip(got, ranges):
            if (a, b) == (0, n):
                assert (g["counts"]["num_cuboids"], g["counts"]["num_leaf_cuboids"]) == PUBLISHED[key]


def test_batch_partial_repeated_and_replayed():
    # batches of different sizes on one core (graphs captured on the second call
    # are replayed by the later ones), with a partition adding up to the totals
    p = sqd.Program.from_lemma("two")
    rng = np.random.default_rng(5)
    with sqd.Core(p) as single, sqd.Core(p, settings=sqd.default_settings(jobs=6, graphs=1)) as batch:
        n, _ = single.shape()
        alone = {}
        for size in (3, 6, 1, 5, 6, 2):
            ranges = _ranges(rng, n, size)
            got = batch.prove_batch(ranges)
            assert len(got) == size
            for (a, b), g in zip(ranges, got):
                if (a, b) not in alone:
                    alone[(a, b)] = single.prove_range(a, b)["counts"]
                assert g["counts"] == alone[(a, b)], (size, a, b)
        cuts = [0, 17, 40, 41, 90, n]
        got = batch.prove_batch(list(zip(cuts[:-1], cuts[1:])))
        tot = sum(np.array([g["counts"]["num_cuboids"], g["counts"]["num_leaf_cuboids"]]) for g in got)
        assert tuple(int(x) for x in tot) == PUBLISHED["two"]
        # single proofs on a batch core run a batch of one
        r = batch.prove()
        assert (r["counts"]["num_cuboids"], r["counts"]["num_leaf_cuboids"]) == PUBLISHED["two"]


@pytest.mark.parametrize("key,fp32", [("ec2", False), ("ec2", True), ("ec3", False)])
def test_batch_criticals_match_oracle(program_text, key, fp32):
    # the systems with expected criticals (ec3: its last static level decided per
    # parent, children that change another variable apart): each job's critical
    # boxes are the oracle's for its range, bit for bit
    p = sqd.Program.from_lemma(key)
    s = sqd.default_settings(generation_count=1, jobs=4, fp32_prefilter=sqd.FP32_ALL_LEVELS if fp32 else 0)
    with sqd.Core(p, settings=s) as core:
        n, _ = core.shape()
        ranges = [(0, n), (0, max(1, n // 3)), (n // 3, n), (n // 2, n // 2 + 1)]
        got = core.prove_batch(ranges, critical_capacity=1 << 19)
    for (a, b), g in zip(ranges, got):
        o = pyoracle.Oracle(program_text(key), fp32=fp32)
        o.set_init_range(a, b)
        oc, ocrit = o.run(generations=1, threads=8, max_criticals=1 << 19)
        assert g["counts"] == oc, (a, b)
        assert g["n_critical"] == oc["num_critical_cuboids"] == len(ocrit)
        nv = g["crit_lb"].shape[0]
        gb = np.vstack([g["crit_lb"], g["crit_ub"]]).T[:, [k for a_ in range(nv) for k in (a_, nv + a_)]]
        ob = ocrit.reshape(len(ocrit), -1)
        gs = gb[np.lexsort(gb.T[::-1])]
        os_ = ob[np.lexsort(ob.T[::-1])]
        assert np.array_equal(gs.view(np.uint64), os_.view(np.uint64)), (a, b)


def test_batch_core_host_calls(program_text):
    # host-driven calls run job 0 alone, before and after batches
    p = sqd.Program.from_lemma("three")
    o = pyoracle.Oracle(program_text("three"))
    qlb, qub, _ = o.initial_queue()
    idx = np.linspace(0, qlb.shape[1] - 1, 64).astype(int)
    lb, ub = qlb[:, idx], qub[:, idx]
    ohl, ohu, ohas, oc = o.nonfinal(lb, ub)
    with sqd.Core(p, settings=sqd.default_settings(jobs=4)) as core:
        for _ in range(2):
            ghl, ghu, ghas, gc = core.nonfinal(lb, ub)
            assert np.array_equal(ghas, ohas) and gc["num_cuboids"] == oc["num_cuboids"]
            keep = ohas.astype(bool)
            assert np.array_equal(ghl[:, keep].view(np.uint64), ohl[:, keep].view(np.uint64))
            n, _ = core.shape()
            got = core.prove_batch([(0, n // 2), (n // 2, n), (0, n)])
            assert got[0]["counts"]["num_cuboids"] + got[1]["counts"]["num_cuboids"] == PUBLISHED["three"][0]
            assert got[2]["counts"]["num_cuboids"] == PUBLISHED["three"][0]


def test_batch_argument_errors():
    p = sqd.Program.from_lemma("top1")
    with sqd.Core(p, settings=sqd.default_settings(jobs=2)) as core:
        n, _ = core.shape()
        with pytest.raises(sqd.SqdError):
            core.prove_batch([(0, 1), (1, 2), (2, 3)])          # more jobs than the core holds
        with pytest.raises(sqd.SqdError):
            core.prove_batch([(0, 1), (5, 5)])                  # empty range
        with pytest.raises(sqd.SqdError):
            core.prove_batch([(0, n + 1)])                      # beyond the subdivisions
        r = core.prove_batch([(0, n)])                          # the core still works
        assert (r[0]["counts"]["num_cuboids"], r[0]["counts"]["num_leaf_cuboids"]) == PUBLISHED["top1"]
    with pytest.raises(sqd.SqdError):
        sqd.Core(p, settings=sqd.default_settings(jobs=65))
