"""The constants that the generated kernels take as literals.  Each program's
binary32 table must hold its binary64 constant intervals converted outward (the
largest float <= each lower bound, the smallest float >= each upper bound, zero
signs included), exactly as the oracle's binary32 conversion gives them; the
oracle's conversion is pinned against exact rationals in test_interval_kat and
the device's against the oracle in test_gpu_fp32.  The binary64 table must be
the program's constant array that the runtime copies to the device."""
import re
from pathlib import Path

import numpy as np
import pytest

import pyoracle

GEN = Path(__file__).resolve().parent.parent / "squares-in-disk_amd" / "gen"
SOURCES = sorted(GEN.glob("*.hip"))


def _literal(s):
    s = s.strip()
    neg = s.startswith("(-")
    if "huge_val" in s:
        return -np.inf if neg else np.inf
    return float.fromhex(s.rstrip("f"))


def _table(src, decl):
    m = re.search(re.escape(decl) + r"\[\] = \{(.*?)\};", src, re.S)
    assert m, decl
    return np.array([_literal(x) for x in m.group(1).split(",")])


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.skipif(not SOURCES, reason="the generated kernel sources are written by the build")
@pytest.mark.parametrize("path", SOURCES, ids=[p.stem for p in SOURCES])
def test_binary32_literals_are_outward(path):
    src = path.read_text()
    k64 = _table(src, "static constexpr double K")
    k32 = _table(src, "static constexpr float K32")
    dev = _table(src, "static const double k_consts")
    assert len(k64) == len(k32) == len(dev) and len(k64) % 2 == 0
    assert np.array_equal(_bits(k64), _bits(dev))
    lo, hi = pyoracle.rounded_op("to32", k64, k64)
    assert np.array_equal(_bits(k32[0::2]), _bits(lo[0::2])), path.stem
    assert np.array_equal(_bits(k32[1::2]), _bits(hi[1::2])), path.stem
    # every pair still encloses its binary64 interval
    assert np.all(k32[0::2] <= k64[0::2]) and np.all(k32[1::2] >= k64[1::2])
