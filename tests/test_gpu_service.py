"""Proof ranges, graph replays, the proof service and the autotuner on the device,
through the C ABI: a range's counts equal the oracle's for the same range, a
partition of the first variable's initial subdivisions adds up to the published
totals, graph replays equal eager runs, and every job's result is the same alone
and in a mix of concurrent jobs."""
import numpy as np
import pytest

import pyoracle
import sqd
from conftest import PUBLISHED

pytestmark = pytest.mark.gpu


def _partition(rng, n, parts):
    cuts = sorted(set(int(x) for x in rng.integers(1, n, parts - 1)))
    edges = [0] + cuts + [n]
    return [(a, b) for a, b in zip(edges[:-1], edges[1:]) if a < b]


@pytest.mark.parametrize("key", ["top1", "one", "two", "three", "five", "six", "four", "more2"])
def test_prove_range_partition_adds_up(key):
    rng = np.random.default_rng(len(key))
    p = sqd.Program.from_lemma(key)
    with sqd.Core(p) as core:
        n, _ = core.shape()
        tot = np.zeros(2, dtype=np.int64)
        for a, b in _partition(rng, n, 5):
            r = core.prove_range(a, b)
            assert r["n_critical"] == 0
            tot += [r["counts"]["num_cuboids"], r["counts"]["num_leaf_cuboids"]]
    assert tuple(int(x) for x in tot) == PUBLISHED[key]


@pytest.mark.parametrize("key,ranges", [("one", [(0, 40), (40, 64), (64, 128), (70, 71)]),
                                        ("three", [(0, 10), (10, 33), (33, 64), (17, 18)]),
                                        ("two", [(90, 128), (5, 6)])])
def test_prove_range_matches_oracle(program_text, key, ranges):
    p = sqd.Program.from_lemma(key)
    for fp32 in (0, sqd.FP32_ALL_LEVELS):
        with sqd.Core(p, settings=sqd.default_settings(fp32_prefilter=fp32)) as core:
            for a, b in ranges:
                g = core.prove_range(a, b)["counts"]
                o = pyoracle.Oracle(program_text(key), fp32=bool(fp32))
                o.set_init_range(a, b)
                oc, _ = o.run(threads=8)
                assert g == oc, (key, a, b, fp32)


def test_graph_replays_match_eager(cores_eager_and_graph):
    # replays of the captured rounds, with other eager work and syncs between them
    eager, graph = cores_eager_and_graph
    n, _ = graph.shape()
    rng = np.random.default_rng(11)
    for rep in range(6):
        a = int(rng.integers(0, n - 1))
        b = int(rng.integers(a + 1, n + 1))
        want = eager.prove_range(a, b)["counts"]
        got = graph.prove_range(a, b)["counts"]
        assert got == want, (rep, a, b)
        full = graph.prove()["counts"]
        assert (full["num_cuboids"], full["num_leaf_cuboids"]) == PUBLISHED["two"]
        qlb, qub, _ = graph.initial_queue()                      # eager work on the same core
        _, _, has, _ = graph.nonfinal(qlb[:, :8], qub[:, :8])
        assert has.shape == (8,)


@pytest.fixture
def cores_eager_and_graph():
    p = sqd.Program.from_lemma("two")
    e = sqd.Core(p, settings=sqd.default_settings(graphs=0))
    g = sqd.Core(p, settings=sqd.default_settings(graphs=1))
    yield e, g
    e.close()
    g.close()


# a mix of jobs: (lemma, first, last)
MIX = [("top1", 0, 128), ("top2", 10, 90), ("one", 0, 128), ("one", 60, 70), ("two", 0, 64), ("two", 64, 128),
       ("three", 0, 64), ("three", 20, 30), ("five", 0, 128), ("six", 0, 64), ("seven", 0, 128), ("four", 30, 34),
       ("more2", 60, 61), ("one", 64, 96), ("two", 100, 101), ("six", 64, 128)]


def test_service_jobs_same_alone_and_in_mix():
    progs = {k: sqd.Program.from_lemma(k) for k in set(j[0] for j in MIX)}
    alone = []
    with sqd.Service(workers=1, fp32_prefilter=0) as svc:            # one job at a time, binary64
        for k, a, b in MIX:
            alone.append(svc.wait(svc.submit(progs[k], a, b)))
    with sqd.Service(workers=8) as svc:                               # all in flight, FP32 prefilter
        ids = [svc.submit(progs[k], a, b) for k, a, b in MIX]
        mixed = [svc.wait(i) for i in ids]
    for (k, a, b), r1, r2 in zip(MIX, alone, mixed):
        assert r1["status"] == 0 and r2["status"] == 0
        assert r1["counts"] == r2["counts"], (k, a, b)
        assert r2["n_critical"] == 0
        assert r2["latency_seconds"] >= r2["run_seconds"] > 0
    # whole statements add up to the published totals
    one = [r["counts"]["num_cuboids"] for (k, a, b), r in zip(MIX, mixed) if k == "two" and (a, b) in ((0, 64), (64, 128))]
    assert sum(one) == PUBLISHED["two"][0]


def test_service_memory_budget_evicts():
    # a budget for about one core: programs take turns, results stay exact
    initial = {"one": 128, "three": 64, "five": 128}            # first variable's initial subdivisions
    progs = {k: sqd.Program.from_lemma(k) for k in initial}
    with sqd.Service(workers=3, memory_budget=1 << 30, level_bytes=64 << 20, frontier_capacity=1 << 20) as svc:
        ids = [(k, svc.submit(progs[k], 0, initial[k])) for k in initial for _ in range(2)]
        for k, j in ids:
            r = svc.wait(j)
            assert (r["counts"]["num_cuboids"], r["counts"]["num_leaf_cuboids"]) == PUBLISHED[k]


def test_autotune_keeps_counts():
    p = sqd.Program.from_lemma("three")
    best, rep = sqd.autotune(p, target_seconds=0.002, repetitions=2)
    assert rep["trials"] >= 2 and rep["best_seconds"] <= rep["plain_seconds"]
    with sqd.Core(p, settings=best) as core:
        r = core.prove_range(rep["first"], rep["last"])
        assert r["counts"] == rep["counts"]
        full = core.prove()["counts"]
        assert (full["num_cuboids"], full["num_leaf_cuboids"]) == PUBLISHED["three"]


def test_progress_reports():
    # progress between rounds and once at the end (ProgressInfo of the reference)
    p = sqd.Program.from_lemma("more1")                    # about 160 rounds
    seen = []
    with sqd.Core(p) as core:
        core.set_progress(seen.append, start_after=0.0, every=0.0)
        r = core.prove()
        core.set_progress(None)
        r2 = core.prove()
    assert r2["counts"] == r["counts"]
    assert len(seen) >= 3 and seen[-1]["done"] and seen[-1]["status"] == 0
    assert all(not s["done"] for s in seen[:-1])
    assert (seen[-1]["cuboid_count"], seen[-1]["leaf_count"]) == PUBLISHED["more1"] and seen[-1]["queue_size"] == 0
    cub = [s["cuboid_count"] for s in seen]
    assert cub == sorted(cub) and [s["rounds"] for s in seen] == sorted(s["rounds"] for s in seen)
    for s in seen[:-1]:
        lb, ub = s["head"]
        assert s["queue_size"] > 0 and all(a <= b for a, b in zip(lb, ub))


def test_service_fixed_settings_and_batches():
    # settings fixed before the first job are the ones the cores use; batches of
    # one program's queued jobs give each job its range's counts
    p = sqd.Program.from_lemma("two")
    s = sqd.default_settings(fp32_prefilter=2, prefilter_group=4, dequeue_buffer_size=4096)
    ranges = [(0, 128), (0, 64), (64, 128), (5, 6), (0, 128), (100, 128), (1, 127)]
    with sqd.Core(p) as core:
        want = [core.prove_range(a, b)["counts"] for a, b in ranges]
    with sqd.Service(workers=2, batch_jobs=3) as svc:
        assert svc.settings(p) is None
        svc.set_settings(p, s)
        got = svc.settings(p)
        assert (got["fp32_prefilter"], got["prefilter_group"], got["dequeue_buffer_size"], got["jobs"]) == (2, 4, 4096, 3)
        ids = [svc.submit(p, a, b) for a, b in ranges]
        res = [svc.wait(i) for i in ids]
        with pytest.raises(sqd.SqdError):
            svc.set_settings(p, s)                           # the program has cores now
    assert [r["counts"] for r in res] == want


def test_service_submit_many():
    # jobs handed over in one call: each gets its range's counts, and a call
    # with an empty range is refused before any of its jobs is queued
    p = sqd.Program.from_lemma("two")
    ranges = [(0, 128), (0, 64), (64, 128), (5, 6), (100, 128), (1, 127), (0, 128), (31, 97)]
    with sqd.Core(p) as core:
        want = [core.prove_range(a, b)["counts"] for a, b in ranges]
    with sqd.Service(workers=2, batch_jobs=4) as svc:
        with pytest.raises(sqd.SqdError):
            svc.submit_many(p, [(0, 8), (9, 9), (10, 20)])
        assert svc.submit_many(p, []) == []
        ids = svc.submit_many(p, ranges)
        assert len(set(ids)) == len(ranges)
        res = [svc.wait(i) for i in ids]
        st = svc.stats()
    assert [r["status"] for r in res] == [0] * len(ranges)
    assert [r["counts"] for r in res] == want
    assert st["jobs_completed"] == len(ranges)          # nothing of the refused call ran
