"""sqd_prove_range_noalloc against sqd_prove_range, called through the C ABI.

The entry point runs one job in a static device array instead of allocated
memory; its results must be exactly those of sqd_prove_range on a core with the
same settings, on two calls in a row (the second on a smaller range, so that
whatever the first left in the array must not leak into it).
"""
import ctypes

import numpy as np
import pytest

import sqd

pytestmark = pytest.mark.gpu


def _noalloc(prog, settings, first, last, capacity=0):
    L = sqd.lib()
    c = sqd.Counts()
    k = ctypes.c_size_t()
    cl = np.zeros((prog.nargs, max(capacity, 1)))
    cu = np.zeros((prog.nargs, max(capacity, 1)))
    rc = L.sqd_prove_range_noalloc(0, prog._h, ctypes.byref(settings), first, last, ctypes.byref(c),
                                   sqd._dptr(cl) if capacity else None, sqd._dptr(cu) if capacity else None,
                                   capacity, ctypes.byref(k))
    if rc != 0:
        raise sqd.SqdError(rc, L.sqd_last_error().decode())
    m = min(k.value, capacity)
    return c.as_dict(), int(k.value), cl[:, :m], cu[:, :m]


def _with_core(prog, settings, first, last, capacity=0):
    L = sqd.lib()
    with sqd.Core(prog, settings=settings) as core:
        c = sqd.Counts()
        k = ctypes.c_size_t()
        cl = np.zeros((prog.nargs, max(capacity, 1)))
        cu = np.zeros((prog.nargs, max(capacity, 1)))
        rc = L.sqd_prove_range(core._h, first, last, ctypes.byref(c), sqd._dptr(cl) if capacity else None,
                               sqd._dptr(cu) if capacity else None, capacity, ctypes.byref(k), None)
        if rc != 0:
            raise sqd.SqdError(rc, L.sqd_last_error().decode())
    m = min(k.value, capacity)
    return c.as_dict(), int(k.value), cl[:, :m], cu[:, :m]


def _sorted(lb, ub):
    m = np.vstack([lb.view(np.uint64), ub.view(np.uint64)]).T
    return m[np.lexsort(m.T[::-1])]


@pytest.mark.parametrize("key,ranges,gens", [("one", [(0, 128), (40, 72)], 8),
                                             ("two", [(0, 128), (100, 101)], 8),
                                             ("ec1", [(0, 8), (2, 5)], 2)])
def test_noalloc_matches_prove_range(key, ranges, gens):
    prog = sqd.Program.from_lemma(key)
    s = sqd.default_settings(generation_count=gens, level_bytes=8 << 20, frontier_capacity=1 << 16,
                             critical_capacity=4096)
    for first, last in ranges:                    # two calls in a row, the second smaller
        got = _noalloc(prog, s, first, last, capacity=4096)
        want = _with_core(prog, s, first, last, capacity=4096)
        assert got[0] == want[0] and got[1] == want[1], (key, first, last, got[:2], want[:2])
        assert np.array_equal(_sorted(got[2], got[3]), _sorted(want[2], want[3]))


def test_noalloc_prefilter_matches_prove_range():
    prog = sqd.Program.from_lemma("three")
    s = sqd.default_settings(level_bytes=4 << 20, frontier_capacity=1 << 18, critical_capacity=4096,
                             fp32_prefilter=sqd.FP32_ALL_LEVELS, prefilter_group=8)
    for first, last in [(0, 64), (10, 12)]:
        got = _noalloc(prog, s, first, last)
        want = _with_core(prog, s, first, last)
        assert got[:2] == want[:2], (first, last, got[:2], want[:2])


def test_noalloc_rejects_settings_beyond_its_array():
    prog = sqd.Program.from_lemma("one")
    s = sqd.default_settings(level_bytes=1 << 30)  # two level buffers of 1 GiB do not fit 64 MiB
    with pytest.raises(sqd.SqdError) as e:
        _noalloc(prog, s, 0, 128)
    assert e.value.code == -4 and "scratch" in str(e.value)
