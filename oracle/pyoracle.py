"""ctypes wrapper of the CPU oracle (test infrastructure only).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker and the CPU baseline.  The product (libsqd.so, the sqd package) never
imports this module.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
CLI = HERE / "_build" / "sqd_oracle"

_lib = None


class OracleCounts(ctypes.Structure):
    _fields_ = [("num_cuboids", ctypes.c_uint64), ("num_leaf_cuboids", ctypes.c_uint64),
                ("num_critical_cuboids", ctypes.c_uint64), ("num_repeated_nodes", ctypes.c_uint64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


_D = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.c_uint64
_P = ctypes.c_void_p


def ensure_built():
    if not LIB.exists() or not CLI.exists():
        subprocess.run(["make", "-C", str(HERE), "-j8"], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        ensure_built()
        L = ctypes.CDLL(str(LIB))
        sig = {
            "oracle_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
            "oracle_free": (None, [_P]),
            "oracle_nargs": (ctypes.c_int, [_P]),
            "oracle_last_error": (ctypes.c_char_p, []),
            "oracle_initial_box": (ctypes.c_int, [_P, _D, _D]),
            "oracle_run": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OracleCounts), _D, _U64,
                                          ctypes.POINTER(_U64)]),
            "oracle_run_slice": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _U64, _U64,
                                                ctypes.POINTER(OracleCounts), ctypes.POINTER(_U64)]),
            "oracle_initial_queue": (ctypes.c_int, [_P, _D, _D, _U64, ctypes.POINTER(_U64), ctypes.POINTER(OracleCounts)]),
            "oracle_nonfinal": (ctypes.c_int, [_P, _U64, _D, _D, _D, _D, ctypes.POINTER(ctypes.c_uint8),
                                               ctypes.POINTER(OracleCounts)]),
            "oracle_final": (ctypes.c_int, [_P, _U64, _D, _D, _D, _D, _U64, ctypes.POINTER(_U64),
                                            ctypes.POINTER(OracleCounts)]),
            "oracle_split": (ctypes.c_int, [_P, _U64, _D, _D, _D, _D, _U64, ctypes.POINTER(_U64),
                                            ctypes.POINTER(OracleCounts)]),
            "oracle_rounded_op": (ctypes.c_int, [ctypes.c_int, _U64, _D, _D, _D, _D]),
            "oracle_flop_count": (_U64, []),
            "oracle_flop_reset": (None, []),
            "oracle_flop_enable": (None, [ctypes.c_int]),
            "oracle_interval_op": (ctypes.c_int, [ctypes.c_int, _U64, _D, _D, _D]),
            "oracle_phase_counts": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
            "oracle_phase_counts32": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_U64)]),
            "oracle_set_fp32_prefilter": (ctypes.c_int, [_P, ctypes.c_int]),
            "oracle_set_init_range": (ctypes.c_int, [_P, _U64, _U64]),
            "oracle_fp32_stats": (ctypes.c_int, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
            "oracle_eval_program": (ctypes.c_int, [_P, ctypes.c_int, _U64, _D, _D, _D, ctypes.POINTER(ctypes.c_uint8)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().oracle_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_D)


class Oracle:
    def __init__(self, program_text, fp32=False):
        self._h = _P()
        _check(lib().oracle_load(program_text.encode(), ctypes.byref(self._h)))
        self.N = lib().oracle_nargs(self._h)
        if fp32:
            _check(lib().oracle_set_fp32_prefilter(self._h, 1))

    def set_init_range(self, first, last):
        """Restrict run() to the initial subdivisions [first, last) of the first dynamic variable."""
        _check(lib().oracle_set_init_range(self._h, first, last))

    def fp32_stats(self):
        """(checked, decided) static-level children of the last call under the FP32 prefilter."""
        a, b = _U64(), _U64()
        _check(lib().oracle_fp32_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def close(self):
        if self._h:
            lib().oracle_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initial_box(self):
        lb = np.zeros(self.N)
        ub = np.zeros(self.N)
        _check(lib().oracle_initial_box(self._h, _dp(lb), _dp(ub)))
        return lb, ub

    def run(self, generations=8, threads=8, max_criticals=0):
        c = OracleCounts()
        crit = np.zeros(max(1, max_criticals) * self.N * 2)
        n = _U64()
        _check(lib().oracle_run(self._h, generations, threads, ctypes.byref(c), _dp(crit), max_criticals,
                                ctypes.byref(n)))
        k = min(n.value, max_criticals)
        boxes = crit[: k * self.N * 2].reshape(k, self.N, 2)
        return c.as_dict(), boxes

    def run_slice(self, first, last, generations=8, threads=8):
        c = OracleCounts()
        iql = _U64()
        _check(lib().oracle_run_slice(self._h, generations, threads, first, last, ctypes.byref(c), ctypes.byref(iql)))
        return c.as_dict(), int(iql.value)

    def initial_queue(self, cap=1 << 20):
        lb = np.zeros(cap * self.N)
        ub = np.zeros(cap * self.N)
        n = _U64()
        c = OracleCounts()
        _check(lib().oracle_initial_queue(self._h, _dp(lb), _dp(ub), cap, ctypes.byref(n), ctypes.byref(c)))
        k = min(n.value, cap)
        return lb[: k * self.N].reshape(k, self.N).T.copy(), ub[: k * self.N].reshape(k, self.N).T.copy(), c.as_dict()

    def eval_program(self, lb, ub, n_bounds, n_constraints, precision=64):
        """Bound values (n, n_bounds, 3) and constraint IBools (n, n_constraints, 2)
        on SoA boxes (N, n), as sqd_eval_program gives them."""
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        b = np.zeros((n, n_bounds, 3))
        c = np.zeros((n, n_constraints, 2), dtype=np.uint8)
        _check(lib().oracle_eval_program(self._h, precision, n, _dp(lb), _dp(ub), _dp(b),
                                         c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return b, c

    def nonfinal(self, lb, ub):
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        hl = np.zeros_like(lb)
        hu = np.zeros_like(ub)
        has = np.zeros(n, dtype=np.uint8)
        c = OracleCounts()
        _check(lib().oracle_nonfinal(self._h, n, _dp(lb), _dp(ub), _dp(hl), _dp(hu),
                                     has.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(c)))
        return hl, hu, has, c.as_dict()

    def final(self, lb, ub, cap=1 << 20):
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        cl = np.zeros(cap * self.N)
        cu = np.zeros(cap * self.N)
        k = _U64()
        c = OracleCounts()
        _check(lib().oracle_final(self._h, n, _dp(lb), _dp(ub), _dp(cl), _dp(cu), cap, ctypes.byref(k), ctypes.byref(c)))
        m = min(k.value, cap)
        return cl[: m * self.N].reshape(m, self.N).T.copy(), cu[: m * self.N].reshape(m, self.N).T.copy(), c.as_dict()

    def split(self, lb, ub, cap=1 << 20):
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        ol = np.zeros(cap * self.N)
        ou = np.zeros(cap * self.N)
        k = _U64()
        c = OracleCounts()
        _check(lib().oracle_split(self._h, n, _dp(lb), _dp(ub), _dp(ol), _dp(ou), cap, ctypes.byref(k), ctypes.byref(c)))
        m = min(k.value, cap)
        return ol[: m * self.N].reshape(m, self.N).T.copy(), ou[: m * self.N].reshape(m, self.N).T.copy(), c.as_dict()


def rounded_op(op, a, b):
    code = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "acos": 5, "acos_series": 6, "add32": 8, "sub32": 9, "mul32": 10,
            "div32": 11, "sqrt32": 12, "acos32": 13, "to32": 14}[op]
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lo = np.zeros_like(a)
    hi = np.zeros_like(a)
    _check(lib().oracle_rounded_op(code, a.size, _dp(a), _dp(b), _dp(lo), _dp(hi)))
    return lo, hi


INTERVAL_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "square": 5, "neg": 6, "max": 7, "min": 8,
                "split": 9, "acos": 10}


def interval_op(op, a, b=None):
    """Interval op on arrays of shape (n, 3) = (lb, ub, undefined); returns (n, 3)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.zeros_like(a) if b is None else np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(a)
    _check(lib().oracle_interval_op(INTERVAL_OPS[op], a.shape[0], _dp(a), _dp(b), _dp(out)))
    return out


def run_cli(program_path, threads=8, generations=8, flops=False, slice_range=None, timeout=None, fp32=False,
            init_range=None):
    """Run the oracle CLI on a program file; returns the parsed JSON line.

    slice_range: initial-queue entries [first, last); init_range: the first
    dynamic variable's initial subdivisions [first, last) (as sqd_prove_range)."""
    import json
    ensure_built()
    cmd = [str(CLI), str(program_path), "--threads", str(threads), "--generations", str(generations)]
    if flops:
        cmd.append("--flops")
    if fp32:
        cmd.append("--fp32")
    if slice_range is not None:
        cmd += ["--slice", str(slice_range[0]), str(slice_range[1])]
    if init_range is not None:
        cmd += ["--range", str(init_range[0]), str(init_range[1])]
    r = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=timeout)
    return json.loads(r.stdout.strip().splitlines()[-1])
