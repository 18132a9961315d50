"""The device's directed binary64 acos (squares-in-disk_amd/csrc/acos_fixed.hpp:
the cos walk from the math library's acos, its second step from the slope, the
series fallback) built for the CPU and compared bit for bit with MPFR's
directed roundings through the oracle (the reference rounds acos with MPFR,
ivarp/src/interval_acos.cpp:29-66), in both rounding modes of the environment.
No GPU needed."""
import ctypes
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle
from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def host_acos():
    if not shutil.which(HIPCC):
        pytest.skip("hipcc not installed")
    out = ROOT / "tests" / "_build"
    out.mkdir(exist_ok=True)
    so = out / "acos_host.so"
    subprocess.run([HIPCC, "-x", "c++", "-O2", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I", str(ROOT / "squares-in-disk_amd" / "csrc"),
                    str(ROOT / "tests" / "acos_host.cpp"),
                    "-o", str(so)], check=True, capture_output=True)
    lib = ctypes.CDLL(str(so))
    lib.acos_host_directed.restype = ctypes.c_int
    lib.acos_host_directed.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_long, ctypes.c_int]

    def run(x, round_down):
        x = np.ascontiguousarray(x, dtype=np.float64)
        lo, hi = np.empty_like(x), np.empty_like(x)
        hard = lib.acos_host_directed(x.ctypes.data, lo.ctypes.data, hi.ctypes.data, x.size, round_down)
        return lo, hi, hard
    return run


def _inputs(seed, n):
    g = np.random.default_rng(seed)
    parts = [
        g.uniform(-1.0, 1.0, n),
        np.cos(g.uniform(0.0, np.pi, n)),                                          # y spread over [0, pi]
        np.cos(np.ldexp(1.0, -g.integers(1, 14, n)) * g.uniform(0.5, 1.5, n)),     # y near 0 ...
        -np.cos(np.ldexp(1.0, -g.integers(1, 14, n)) * g.uniform(0.5, 1.5, n)),    # ... and near pi
        np.ldexp(g.uniform(-1.0, 1.0, n), -g.integers(0, 60, n)),                  # x near 0
        (1.0 - np.ldexp(1.0, -g.integers(1, 53, n))) * g.choice([-1.0, 1.0], n),   # x near +-1
    ]
    x = np.concatenate(parts)
    return np.concatenate([x, np.nextafter(x, 2.0), np.nextafter(x, -2.0), [-1.0, 1.0, 0.0, -0.0]])


@pytest.mark.parametrize("round_down", [0, 1])
def test_host_acos_bit_exact_with_mpfr(host_acos, round_down):
    x = _inputs(23 + round_down, 1 << 14)
    x = x[(x >= -1.0) & (x <= 1.0)]
    lo, hi, hard = host_acos(x, round_down)
    olo, ohi = pyoracle.rounded_op("acos", x, x)
    assert hard == 0
    assert np.array_equal(lo.view(np.uint64), olo.view(np.uint64))
    assert np.array_equal(hi.view(np.uint64), ohi.view(np.uint64))
