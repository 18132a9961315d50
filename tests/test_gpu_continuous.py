"""Continuous batches: jobs join free rows of a running batch at round
boundaries and leave when their frontier empties; each job's counts and
criticals must be exactly those of its range proved alone (the reference proves
one statement at a time, prover/proof_driver.hpp:393-438, so "alone" is the
reference's semantics: the same counts as the full proof's subtrees)."""
import time

import numpy as np
import pytest

import sqd
from conftest import PUBLISHED

pytestmark = pytest.mark.gpu


def _alone(prog, ranges, **kw):
    out = {}
    with sqd.Core(prog, settings=sqd.default_settings(**kw)) as core:
        for a, b in ranges:
            out[(a, b)] = core.prove_range(a, b)
    return out


# three: three dynamic variables (five's proof is almost all decided in its
# initial queue and ends in one round)
@pytest.mark.parametrize("key,fp32,dequeue", [("two", False, 256), ("two", True, 256), ("three", False, 256)])
def test_jobs_join_a_running_batch(key, fp32, dequeue):
    prog = sqd.Program.from_lemma(key)
    kw = dict(dequeue_buffer_size=dequeue, fp32_prefilter=sqd.FP32_ALL_LEVELS if fp32 else 0)
    with sqd.Core(prog, settings=sqd.default_settings(**kw)) as c:
        n, _ = c.shape()
    first = (0, n)                                           # a long job in row 0
    later = [(n // 4, n // 4 + 3), (n // 2, n // 2 + 1), (3 * n // 4, n)]
    want = _alone(prog, [first] + later, **kw)
    with sqd.Core(prog, settings=sqd.default_settings(jobs=4, **kw)) as core:
        core.batch_admit(0, *first)
        done, active = core.batch_step(4)
        assert active == 1 and not done                      # the long job runs on
        for r, rg in enumerate(later, start=1):              # the others join it
            core.batch_admit(r, *rg)
        got, steps = {}, 0
        while True:
            done, active = core.batch_step(4)
            steps += 1
            got.update(done)
            if active == 0:
                break
            assert steps < 100000
        assert set(got) == {0, 1, 2, 3}
        for r, rg in enumerate([first] + later):
            assert got[r]["counts"] == want[rg]["counts"], (key, fp32, r, rg)
        if key == "two" and not fp32:
            assert got[0]["counts"]["num_cuboids"] == PUBLISHED["two"][0]
        # a row can take the next job after its job ended
        core.batch_admit(1, *later[1])
        done, active = core.batch_step(4)
        while active:
            more, active = core.batch_step(4)
            done.update(more)
        assert done[1]["counts"] == want[later[1]]["counts"]


def test_continuous_batch_criticals_match_alone():
    # the 2-variable known-answer system (ivarp/test/expected_criticals.cpp): 128 criticals
    prog = sqd.Program.from_lemma("ec1")
    kw = dict(generation_count=2)
    with sqd.Core(prog, settings=sqd.default_settings(**kw)) as core:
        n, _ = core.shape()
        alone = core.prove()
    with sqd.Core(prog, settings=sqd.default_settings(jobs=2, **kw)) as core:
        core.batch_admit(1, 0, n)
        done, active = core.batch_step(2)
        while active:
            more, active = core.batch_step(2)
            done.update(more)
        assert done[1]["counts"] == alone["counts"] and done[1]["n_critical"] == 128
        cl, cu = np.zeros((2, 128)), np.zeros((2, 128))
        k = sqd._SZ()
        sqd._check(sqd.lib().sqd_core_job_criticals(core._h, 1, sqd._dptr(cl), sqd._dptr(cu), 128, sqd.ctypes.byref(k)))
        got = np.vstack([cl, cu]).T.view(np.uint64)
        ref = np.vstack([alone["crit_lb"], alone["crit_ub"]]).T.view(np.uint64)
        assert np.array_equal(got[np.lexsort(got.T[::-1])], ref[np.lexsort(ref.T[::-1])])


def test_service_admits_jobs_into_a_running_batch():
    # one worker: a long more2 job starts a batch; jobs submitted while it runs
    # join that batch at a round boundary instead of waiting for it
    prog = sqd.Program.from_lemma("more2")
    n = 128
    ranges = [(8 * i, 8 * i + 2) for i in range(1, 6)]
    want = _alone(prog, [(0, n)] + ranges, fp32_prefilter=0, dequeue_buffer_size=2048)
    with sqd.Service(workers=1, batch_jobs=8, fp32_prefilter=0) as svc:
        s = sqd.default_settings(fp32_prefilter=0, dequeue_buffer_size=2048)
        svc.set_settings(prog, s)
        warm = svc.submit(prog, 0, 1)
        svc.wait(warm)
        long_job = svc.submit(prog, 0, n)
        time.sleep(0.01)
        ids = [svc.submit(prog, a, b) for a, b in ranges]
        res = [svc.wait(i) for i in ids]
        first = svc.wait(long_job)
    assert first["counts"] == want[(0, n)]["counts"]
    assert first["counts"]["num_cuboids"] == PUBLISHED["more2"][0]
    for rg, r in zip(ranges, res):
        assert r["status"] == 0 and r["counts"] == want[rg]["counts"], rg
    assert any(r["joined_running"] for r in res)
    assert all(r["latency_seconds"] < first["latency_seconds"] for r in res if r["joined_running"])


def test_service_admission_by_frontier_bytes():
    # a frontier budget limits the jobs that run together; the counts stay the
    # counts of each range proved alone
    prog = sqd.Program.from_lemma("two")
    ranges = [(16 * i, 16 * i + 16) for i in range(8)]
    kw = dict(fp32_prefilter=0, dequeue_buffer_size=256)
    want = _alone(prog, ranges, **kw)
    peaks = {}
    for budget in (0, 1):
        with sqd.Service(workers=2, batch_jobs=8, fp32_prefilter=0, frontier_budget=budget) as svc:
            svc.set_settings(prog, sqd.default_settings(**kw))
            ids = [svc.submit(prog, a, b) for a, b in ranges]
            res = [svc.wait(i) for i in ids]
            st = svc.stats()
        for rg, r in zip(ranges, res):
            assert r["status"] == 0 and r["counts"] == want[rg]["counts"], (budget, rg)
        assert st["jobs_completed"] == len(ranges) and st["rows_running"] == 0 and st["frontier_bytes"] == 0
        peaks[budget] = st["rows_peak"]
    assert peaks[0] > 1        # without a budget the batch takes several jobs at once
    assert peaks[1] == 1       # a budget of one byte: a job is admitted only when none runs
