"""Python binding of the MI355X prover core (libsqd.so, C ABI in include/sqd.h).

This is host plumbing over the C ABI; every computation runs in the HIP
kernels of libsqd.so.  Loading fails loudly when the library has not been built.

    prog = Program.from_lemma("one")
    with Core(prog) as core:
        result = core.prove()          # counts, criticals, seconds
"""
import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libsqd.so"

_lib = None


class SqdError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("sqd error %d: %s" % (code, message))
        self.code = code


class Counts(ctypes.Structure):
    _fields_ = [("num_cuboids", ctypes.c_uint64), ("num_leaf_cuboids", ctypes.c_uint64),
                ("num_critical_cuboids", ctypes.c_uint64), ("num_repeated_nodes", ctypes.c_uint64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class Settings(ctypes.Structure):
    _fields_ = [("generation_count", ctypes.c_uint32), ("dequeue_buffer_size", ctypes.c_uint32),
                ("frontier_capacity", ctypes.c_uint64), ("level_bytes", ctypes.c_uint64),
                ("critical_capacity", ctypes.c_uint64), ("grid_blocks", ctypes.c_uint32),
                ("fp32_prefilter", ctypes.c_uint32), ("graphs", ctypes.c_uint32), ("prefilter_group", ctypes.c_uint32),
                ("jobs", ctypes.c_uint32), ("group_levels", ctypes.c_uint32)]


FP32_ALL_LEVELS = 0xFFFFFFFF


class WidthStat(ctypes.Structure):
    _fields_ = [("constraint", ctypes.c_int32), ("reserved", ctypes.c_uint32), ("samples", ctypes.c_uint64),
                ("mean_ratio", ctypes.c_double), ("max_ratio", ctypes.c_double)]


class AutotuneOptions(ctypes.Structure):
    _fields_ = [("target_seconds", ctypes.c_double), ("repetitions", ctypes.c_int32), ("fp32_first", ctypes.c_uint32)]


class AutotuneReport(ctypes.Structure):
    _fields_ = [("first", ctypes.c_uint32), ("last", ctypes.c_uint32), ("trials", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("plain_seconds", ctypes.c_double), ("best_seconds", ctypes.c_double),
                ("counts", Counts)]


class ServiceOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("workers", ctypes.c_uint32), ("memory_budget", ctypes.c_uint64),
                ("fp32_prefilter", ctypes.c_uint32), ("autotune", ctypes.c_uint32), ("level_bytes", ctypes.c_uint64),
                ("frontier_capacity", ctypes.c_uint64), ("core_bytes_estimate", ctypes.c_uint64),
                ("batch_jobs", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("frontier_budget", ctypes.c_uint64)]


class ServiceStats(ctypes.Structure):
    _fields_ = [("frontier_bytes", ctypes.c_uint64), ("frontier_bytes_peak", ctypes.c_uint64),
                ("rows_running", ctypes.c_uint32), ("rows_peak", ctypes.c_uint32),
                ("cores_created", ctypes.c_uint64), ("cores_destroyed", ctypes.c_uint64),
                ("jobs_completed", ctypes.c_uint64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class JobCounts(ctypes.Structure):
    _fields_ = [("counts", Counts), ("n_critical", ctypes.c_uint64), ("fp32_checked", ctypes.c_uint64),
                ("fp32_decided", ctypes.c_uint64), ("frontier_peak", ctypes.c_uint64)]


class JobResult(ctypes.Structure):
    _fields_ = [("counts", Counts), ("n_critical", ctypes.c_uint64), ("queued_seconds", ctypes.c_double),
                ("run_seconds", ctypes.c_double), ("latency_seconds", ctypes.c_double), ("status", ctypes.c_int32),
                ("joined_running", ctypes.c_uint32), ("frontier_peak", ctypes.c_uint64)]

    def as_dict(self):
        return {"counts": self.counts.as_dict(), "n_critical": int(self.n_critical),
                "queued_seconds": self.queued_seconds, "run_seconds": self.run_seconds,
                "latency_seconds": self.latency_seconds, "status": int(self.status),
                "joined_running": bool(self.joined_running), "frontier_peak": int(self.frontier_peak)}


class Progress(ctypes.Structure):
    _fields_ = [("queue_size", ctypes.c_uint64), ("cuboid_count", ctypes.c_uint64), ("leaf_count", ctypes.c_uint64),
                ("critical_count", ctypes.c_uint64), ("rounds", ctypes.c_uint64), ("done", ctypes.c_int32),
                ("status", ctypes.c_int32), ("head_lb", ctypes.POINTER(ctypes.c_double)),
                ("head_ub", ctypes.POINTER(ctypes.c_double))]


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(Progress))
CRITICAL_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.c_size_t, ctypes.c_size_t)


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("mode", ctypes.c_int32), ("level", ctypes.c_int32),
                ("launches", ctypes.c_uint64), ("milliseconds", ctypes.c_double), ("children", ctypes.c_uint64)]

    def as_dict(self):
        return {"name": self.name.decode(), "mode": self.mode, "level": self.level, "launches": int(self.launches),
                "milliseconds": self.milliseconds, "children": int(self.children)}


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_SZ = ctypes.c_size_t


def lib():
    """Load libsqd.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SqdError(-2, "libsqd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(str(LIB_PATH))
        sig = {
            "sqd_settings_default": (None, [ctypes.POINTER(Settings)]),
            "sqd_program_from_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
            "sqd_program_from_lemma": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
            "sqd_program_destroy": (None, [_P]),
            "sqd_program_text": (ctypes.c_int, [_P, ctypes.c_char_p, _SZ, ctypes.POINTER(_SZ)]),
            "sqd_program_nargs": (ctypes.c_int, [_P]),
            "sqd_program_has_kernels": (ctypes.c_int, [_P]),
            "sqd_lemma_count": (ctypes.c_int, []),
            "sqd_lemma_key": (ctypes.c_char_p, [ctypes.c_int]),
            "sqd_lemma_title": (ctypes.c_char_p, [ctypes.c_int]),
            "sqd_core_create": (ctypes.c_int, [ctypes.c_int, _P, ctypes.POINTER(Settings), ctypes.POINTER(_P)]),
            "sqd_core_destroy": (None, [_P]),
            "sqd_initial_box": (ctypes.c_int, [_P, _D, _D]),
            "sqd_initial_queue": (ctypes.c_int, [_P, _D, _D, _SZ, ctypes.POINTER(_SZ), ctypes.POINTER(Counts)]),
            "sqd_nonfinal": (ctypes.c_int, [_P, _D, _D, _SZ, _D, _D, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(Counts)]),
            "sqd_final": (ctypes.c_int, [_P, _D, _D, _SZ, _D, _D, _SZ, ctypes.POINTER(_SZ), ctypes.POINTER(Counts)]),
            "sqd_split": (ctypes.c_int, [_P, _D, _D, _SZ, _D, _D, _SZ, ctypes.POINTER(_SZ), ctypes.POINTER(Counts)]),
            "sqd_prove": (ctypes.c_int, [_P, ctypes.POINTER(Counts), _D, _D, _SZ, ctypes.POINTER(_SZ),
                                         ctypes.POINTER(ctypes.c_double)]),
            "sqd_core_criticals": (ctypes.c_int, [_P, _D, _D, _SZ, ctypes.POINTER(_SZ)]),
            "sqd_prove_range": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Counts), _D, _D, _SZ,
                                               ctypes.POINTER(_SZ), ctypes.POINTER(ctypes.c_double)]),
            "sqd_prove_batch": (ctypes.c_int, [_P, _SZ, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.POINTER(JobCounts), ctypes.POINTER(ctypes.c_double)]),
            "sqd_core_job_criticals": (ctypes.c_int, [_P, _SZ, _D, _D, _SZ, ctypes.POINTER(_SZ)]),
            "sqd_batch_admit": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
            "sqd_batch_step": (ctypes.c_int, [_P, ctypes.POINTER(JobCounts), ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_uint32)]),
            "sqd_batch_frontier": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), _SZ]),
            "sqd_core_shape": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
            "sqd_core_memory": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
            "sqd_autotune_options_default": (None, [ctypes.POINTER(AutotuneOptions)]),
            "sqd_autotune": (ctypes.c_int, [ctypes.c_int, _P, ctypes.POINTER(Settings), ctypes.POINTER(AutotuneOptions),
                                            ctypes.POINTER(Settings), ctypes.POINTER(AutotuneReport)]),
            "sqd_service_options_default": (None, [ctypes.POINTER(ServiceOptions)]),
            "sqd_service_create": (ctypes.c_int, [ctypes.POINTER(ServiceOptions), ctypes.POINTER(_P)]),
            "sqd_service_submit": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
            "sqd_service_submit_many": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                                       ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)]),
            "sqd_service_poll": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(JobResult)]),
            "sqd_service_get_settings": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Settings)]),
            "sqd_service_set_settings": (ctypes.c_int, [_P, _P, ctypes.POINTER(Settings)]),
            "sqd_service_wait": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(JobResult)]),
            "sqd_service_destroy": (None, [_P]),
            "sqd_service_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(ServiceStats)]),
            "sqd_core_set_profiling": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "sqd_core_kernel_stats": (ctypes.c_int, [_P, ctypes.POINTER(KernelStat), _SZ, ctypes.POINTER(_SZ)]),
            "sqd_rounded_op": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _SZ, _D, _D, _D, _D]),
            "sqd_interval_op": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _SZ, _D, _D, _D]),
            "sqd_test_acos_fallbacks": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
            "sqd_core_acos_fallbacks": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
            "sqd_core_set_progress": (ctypes.c_int, [_P, PROGRESS_FN, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
            "sqd_core_set_critical_sink": (ctypes.c_int, [_P, CRITICAL_FN, ctypes.c_void_p]),
            "sqd_eval_program": (ctypes.c_int, [_P, ctypes.c_int, _D, _D, _SZ, _D, ctypes.POINTER(ctypes.c_uint8),
                                                ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
            "sqd_prove_range_noalloc": (ctypes.c_int, [ctypes.c_int, _P, ctypes.POINTER(Settings), ctypes.c_uint32,
                                                       ctypes.c_uint32, ctypes.POINTER(Counts), _D, _D, _SZ,
                                                       ctypes.POINTER(_SZ)]),
            "sqd_core_fp32_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
            "sqd_width_stats": (ctypes.c_int, [_P, _D, _D, _SZ, ctypes.POINTER(WidthStat), _SZ, ctypes.POINTER(_SZ)]),
            "sqd_last_error": (ctypes.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names of the C ABI functions declared in include/sqd.h."""
    return ["sqd_settings_default", "sqd_program_from_text", "sqd_program_from_lemma", "sqd_program_destroy",
            "sqd_program_text", "sqd_program_nargs", "sqd_program_has_kernels", "sqd_lemma_count",
            "sqd_lemma_key", "sqd_lemma_title", "sqd_core_create", "sqd_core_destroy", "sqd_initial_box",
            "sqd_initial_queue", "sqd_nonfinal", "sqd_final", "sqd_split", "sqd_prove", "sqd_prove_range",
            "sqd_prove_batch", "sqd_core_job_criticals", "sqd_batch_admit", "sqd_batch_step", "sqd_batch_frontier",
            "sqd_core_criticals", "sqd_core_shape", "sqd_core_memory", "sqd_autotune_options_default", "sqd_autotune",
            "sqd_service_options_default", "sqd_service_create", "sqd_service_submit", "sqd_service_submit_many",
            "sqd_service_poll",
            "sqd_service_get_settings", "sqd_service_set_settings",
            "sqd_service_wait", "sqd_service_destroy", "sqd_service_get_stats",
            "sqd_core_set_profiling",
            "sqd_core_kernel_stats", "sqd_core_acos_fallbacks", "sqd_core_fp32_stats", "sqd_width_stats",
            "sqd_core_set_progress", "sqd_core_set_critical_sink", "sqd_prove_range_noalloc", "sqd_eval_program",
            "sqd_rounded_op", "sqd_interval_op",
            "sqd_test_acos_fallbacks", "sqd_last_error"]


def _check(rc):
    if rc != 0:
        raise SqdError(rc, lib().sqd_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(_D)


def lemmas():
    L = lib()
    return [(L.sqd_lemma_key(i).decode(), L.sqd_lemma_title(i).decode()) for i in range(L.sqd_lemma_count())]


def default_settings(**kw):
    s = Settings()
    lib().sqd_settings_default(ctypes.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


class Program:
    """A lowered program bound to its compiled kernels."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_lemma(cls, key):
        h = _P()
        _check(lib().sqd_program_from_lemma(key.encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_text(cls, text):
        h = _P()
        _check(lib().sqd_program_from_text(text.encode(), ctypes.byref(h)))
        return cls(h)

    @property
    def nargs(self):
        return lib().sqd_program_nargs(self._h)

    @property
    def has_kernels(self):
        return bool(lib().sqd_program_has_kernels(self._h))

    def text(self):
        n = _SZ()
        _check(lib().sqd_program_text(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().sqd_program_text(self._h, buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    def close(self):
        if self._h:
            lib().sqd_program_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Core:
    """One prover core on one device (sqd_core)."""

    def __init__(self, program, device=0, settings=None):
        self.program = program
        self.N = program.nargs
        self._crit = None          # the critical sink's callback, while one is registered
        self._h = _P()
        s = settings if settings is not None else default_settings()
        _check(lib().sqd_core_create(device, program._h, ctypes.byref(s), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().sqd_core_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def initial_box(self):
        lb = np.zeros(self.N)
        ub = np.zeros(self.N)
        _check(lib().sqd_initial_box(self._h, _dptr(lb), _dptr(ub)))
        return lb, ub

    def initial_queue(self, capacity=1 << 20):
        lb = np.zeros((self.N, capacity))
        ub = np.zeros((self.N, capacity))
        n = _SZ()
        c = Counts()
        _check(lib().sqd_initial_queue(self._h, _dptr(lb), _dptr(ub), capacity, ctypes.byref(n), ctypes.byref(c)))
        m = min(n.value, capacity)
        return lb[:, :m].copy(), ub[:, :m].copy(), c.as_dict()

    def nonfinal(self, lb, ub):
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        hl = np.zeros_like(lb)
        hu = np.zeros_like(ub)
        has = np.zeros(n, dtype=np.uint8)
        c = Counts()
        _check(lib().sqd_nonfinal(self._h, _dptr(lb), _dptr(ub), n, _dptr(hl), _dptr(hu),
                                   has.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(c)))
        return hl, hu, has, c.as_dict()

    def final(self, lb, ub, capacity=1 << 20):
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        cl = np.zeros((self.N, capacity))
        cu = np.zeros((self.N, capacity))
        k = _SZ()
        c = Counts()
        _check(lib().sqd_final(self._h, _dptr(lb), _dptr(ub), n, _dptr(cl), _dptr(cu), capacity, ctypes.byref(k),
                                ctypes.byref(c)))
        m = 0 if self._crit is not None else min(k.value, capacity)   # a sink received them
        return cl[:, :m].copy(), cu[:, :m].copy(), c.as_dict()

    def split(self, lb, ub, capacity=1 << 20):
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        ol = np.zeros((self.N, capacity))
        ou = np.zeros((self.N, capacity))
        k = _SZ()
        c = Counts()
        _check(lib().sqd_split(self._h, _dptr(lb), _dptr(ub), n, _dptr(ol), _dptr(ou), capacity, ctypes.byref(k),
                                ctypes.byref(c)))
        m = min(k.value, capacity)
        return ol[:, :m].copy(), ou[:, :m].copy(), c.as_dict()

    def prove(self, critical_capacity=1 << 16):
        c = Counts()
        k = _SZ()
        secs = ctypes.c_double()
        _check(lib().sqd_prove(self._h, ctypes.byref(c), None, None, 0, ctypes.byref(k), ctypes.byref(secs)))
        m = 0 if self._crit is not None else min(k.value, critical_capacity)   # a sink received them
        cl = np.zeros((self.N, m))
        cu = np.zeros((self.N, m))
        if m:
            n = _SZ()
            _check(lib().sqd_core_criticals(self._h, _dptr(cl), _dptr(cu), m, ctypes.byref(n)))
        return {"counts": c.as_dict(), "crit_lb": cl, "crit_ub": cu, "n_critical": int(k.value),
                "seconds": secs.value}


    def prove_range(self, first, last):
        """Proof restricted to the first variable's initial subdivisions [first, last)."""
        c = Counts()
        k = _SZ()
        secs = ctypes.c_double()
        _check(lib().sqd_prove_range(self._h, first, last, ctypes.byref(c), None, None, 0, ctypes.byref(k),
                                     ctypes.byref(secs)))
        return {"counts": c.as_dict(), "n_critical": int(k.value), "seconds": secs.value}

    def prove_batch(self, ranges, critical_capacity=0):
        """Jobs [(first, last), ...] of this program in the same launches (at most
        the core's settings.jobs); one dict per job as prove_range gives, plus the
        job's criticals (up to critical_capacity) and FP32 prefilter counts."""
        n = len(ranges)
        first = (ctypes.c_uint32 * max(n, 1))(*[int(a) for a, _ in ranges])
        last = (ctypes.c_uint32 * max(n, 1))(*[int(b) for _, b in ranges])
        res = (JobCounts * max(n, 1))()
        secs = ctypes.c_double()
        _check(lib().sqd_prove_batch(self._h, n, first, last, res, ctypes.byref(secs)))
        out = []
        for j in range(n):
            r = res[j]
            d = {"counts": r.counts.as_dict(), "n_critical": int(r.n_critical), "seconds": secs.value,
                 "fp32_stats": (int(r.fp32_checked), int(r.fp32_decided)), "frontier_peak": int(r.frontier_peak)}
            m = min(int(r.n_critical), critical_capacity)
            if critical_capacity:
                cl = np.zeros((self.N, m))
                cu = np.zeros((self.N, m))
                if m:
                    k = _SZ()
                    _check(lib().sqd_core_job_criticals(self._h, j, _dptr(cl), _dptr(cu), m, ctypes.byref(k)))
                d["crit_lb"], d["crit_ub"] = cl, cu
            out.append(d)
        return out

    def batch_admit(self, row, first, last):
        """Continuous batch: the job [first, last) takes the free row `row` and
        starts with the next batch_step."""
        _check(lib().sqd_batch_admit(self._h, row, first, last))

    def batch_frontier(self, rows):
        """Cuboids on each row's frontier after the last batch step."""
        arr = (ctypes.c_uint64 * rows)()
        _check(lib().sqd_batch_frontier(self._h, arr, rows))
        return [int(v) for v in arr]

    def batch_step(self, rows):
        """One round of the continuous batch on a core of `rows` rows: a dict
        {row: result} of the jobs that ended in it, and the rows still running."""
        res = (JobCounts * rows)()
        fin = ctypes.c_uint64()
        act = ctypes.c_uint32()
        _check(lib().sqd_batch_step(self._h, res, ctypes.byref(fin), ctypes.byref(act)))
        out = {}
        for r in range(rows):
            if (fin.value >> r) & 1:
                out[r] = {"counts": res[r].counts.as_dict(), "n_critical": int(res[r].n_critical),
                          "fp32_stats": (int(res[r].fp32_checked), int(res[r].fp32_decided)),
                          "frontier_peak": int(res[r].frontier_peak)}
        return out, int(act.value)

    def set_progress(self, fn, start_after=0.0, every=0.0):
        """fn(dict) after each due round and at the end of sqd_prove (None: off)."""
        if fn is None:
            self._progress = None
            _check(lib().sqd_core_set_progress(self._h, PROGRESS_FN(), None, 0.0, 0.0))
            return
        n = self.N

        def cb(_user, p):
            q = p.contents
            head = None
            if q.head_lb:
                head = ([q.head_lb[a] for a in range(n)], [q.head_ub[a] for a in range(n)])
            fn({"queue_size": int(q.queue_size), "cuboid_count": int(q.cuboid_count),
                "leaf_count": int(q.leaf_count), "critical_count": int(q.critical_count), "rounds": int(q.rounds),
                "done": bool(q.done), "status": int(q.status), "head": head})
        self._progress = PROGRESS_FN(cb)              # kept alive while registered
        _check(lib().sqd_core_set_progress(self._h, self._progress, None, start_after, every))

    def set_critical_sink(self, fn):
        """fn(lb, ub), arrays of shape (N, n), for every chunk of criticals that
        prove, prove_range and final find (None: off); those calls then return no
        critical boxes themselves."""
        if fn is None:
            self._crit = None
            _check(lib().sqd_core_set_critical_sink(self._h, CRITICAL_FN(), None))
            return
        n_args = self.N

        def cb(_user, lb, ub, n, stride):
            a = np.ctypeslib.as_array(lb, shape=(n_args * stride,)).reshape(n_args, stride)[:, :n].copy()
            b = np.ctypeslib.as_array(ub, shape=(n_args * stride,)).reshape(n_args, stride)[:, :n].copy()
            fn(a, b)
        self._crit = CRITICAL_FN(cb)                  # kept alive while registered
        _check(lib().sqd_core_set_critical_sink(self._h, self._crit, None))

    def shape(self):
        """(initial subdivisions of the first dynamic variable, static levels)."""
        a, b = ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().sqd_core_shape(self._h, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def memory(self):
        n = ctypes.c_uint64()
        _check(lib().sqd_core_memory(self._h, ctypes.byref(n)))
        return int(n.value)

    def acos_fallbacks(self):
        """Undecided acos roundings of this core's program (results widened by one double)."""
        n = ctypes.c_uint64()
        _check(lib().sqd_core_acos_fallbacks(self._h, ctypes.byref(n)))
        return int(n.value)

    def fp32_stats(self):
        """(checked, decided): static-level children of the last call that the FP32
        prefilter evaluated and pruned without binary64."""
        a = ctypes.c_uint64()
        b = ctypes.c_uint64()
        _check(lib().sqd_core_fp32_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def width_stats(self, lb, ub):
        """Per constraint: FP32-versus-FP64 width ratio of its compared difference
        over the boxes (SoA arrays of shape (N, n))."""
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        k = _SZ()
        _check(lib().sqd_width_stats(self._h, _dptr(lb), _dptr(ub), n, None, 0, ctypes.byref(k)))
        arr = (WidthStat * max(1, k.value))()
        _check(lib().sqd_width_stats(self._h, _dptr(lb), _dptr(ub), n, arr, k.value, ctypes.byref(k)))
        return [{"constraint": arr[i].constraint, "samples": int(arr[i].samples), "mean_ratio": arr[i].mean_ratio,
                 "max_ratio": arr[i].max_ratio} for i in range(k.value)]

    def eval_program(self, lb, ub, precision=64):
        """Bound and constraint values on boxes (SoA arrays of shape (N, n)):
        (bounds (n, n_bounds, 3) = lb, ub, undefined; constraints (n, n_cons, 2) =
        definitely, possibly), in binary64 or (precision=32) binary32."""
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        n = lb.shape[1]
        nb, nc = _SZ(), _SZ()
        _check(lib().sqd_eval_program(self._h, precision, _dptr(lb), _dptr(ub), n, None, None, ctypes.byref(nb),
                                      ctypes.byref(nc)))
        b = np.zeros((n, nb.value, 3))
        c = np.zeros((n, nc.value, 2), dtype=np.uint8)
        _check(lib().sqd_eval_program(self._h, precision, _dptr(lb), _dptr(ub), n, _dptr(b),
                                      c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(nb),
                                      ctypes.byref(nc)))
        return b, c

    def set_profiling(self, enable=True, only=None):
        """Time launches with HIP events on the core's stream (clears the statistics).

        only: (mode, level) of the one level kernel to time, or None for every launch."""
        m, l = only if only is not None else (-1, -1)
        _check(lib().sqd_core_set_profiling(self._h, 1 if enable else 0, m, l))

    def kernel_stats(self):
        """Per-kernel-kind launches and summed event milliseconds since profiling was enabled."""
        n = _SZ()
        _check(lib().sqd_core_kernel_stats(self._h, None, 0, ctypes.byref(n)))
        arr = (KernelStat * max(1, n.value))()
        _check(lib().sqd_core_kernel_stats(self._h, arr, n.value, ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]


def rounded_op(op, a, b, device=0):
    """Device RD/RU of a op b (op: 'add', 'sub', 'mul', 'div', 'sqrt')."""
    code = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "acos": 5, "acos_series": 6, "add32": 8, "sub32": 9, "mul32": 10,
            "div32": 11, "sqrt32": 12, "acos32": 13, "to32": 14}[op]
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lo = np.zeros_like(a)
    hi = np.zeros_like(a)
    _check(lib().sqd_rounded_op(device, code, a.size, _dptr(a), _dptr(b), _dptr(lo), _dptr(hi)))
    return lo, hi


def acos_fallbacks(device=0):
    """Undecided acos roundings in the test kernels since the library was loaded."""
    n = ctypes.c_uint64()
    _check(lib().sqd_test_acos_fallbacks(device, ctypes.byref(n)))
    return int(n.value)


INTERVAL_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "square": 5, "neg": 6, "max": 7, "min": 8,
                "split": 9, "acos": 10}


def interval_op(op, a, b=None, device=0):
    """Device interval op on arrays of shape (n, 3) = (lb, ub, undefined); returns (n, 3)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.zeros_like(a) if b is None else np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(a)
    _check(lib().sqd_interval_op(device, INTERVAL_OPS[op], a.shape[0], _dptr(a), _dptr(b), _dptr(out)))
    return out


def autotune(program, device=0, base=None, target_seconds=0.02, repetitions=3, fp32_first=False):
    """sqd_autotune: the best settings found and the report (dict)."""
    o = AutotuneOptions()
    lib().sqd_autotune_options_default(ctypes.byref(o))
    o.target_seconds = target_seconds
    o.repetitions = repetitions
    o.fp32_first = 1 if fp32_first else 0
    best = Settings()
    rep = AutotuneReport()
    _check(lib().sqd_autotune(device, program._h, ctypes.byref(base) if base is not None else None, ctypes.byref(o),
                              ctypes.byref(best), ctypes.byref(rep)))
    return best, {"first": rep.first, "last": rep.last, "trials": rep.trials, "plain_seconds": rep.plain_seconds,
                  "best_seconds": rep.best_seconds, "counts": rep.counts.as_dict(),
                  "settings": {f: getattr(best, f) for f, _ in Settings._fields_}}


class Service:
    """The proof service (sqd_service): submit jobs, wait for their results."""

    def __init__(self, device=0, **options):
        o = ServiceOptions()
        lib().sqd_service_options_default(ctypes.byref(o))
        o.device = device
        for k, v in options.items():
            setattr(o, k, v)
        self._h = _P()
        _check(lib().sqd_service_create(ctypes.byref(o), ctypes.byref(self._h)))

    def settings(self, program):
        """The settings of the program's cores (dict), or None before its first job."""
        known = ctypes.c_int()
        s = Settings()
        _check(lib().sqd_service_get_settings(self._h, program._h, ctypes.byref(known), ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Settings._fields_} if known.value else None

    def set_settings(self, program, settings):
        """Fixes the program's settings (a Settings) before its first job."""
        _check(lib().sqd_service_set_settings(self._h, program._h, ctypes.byref(settings)))

    def submit(self, program, first, last):
        j = ctypes.c_uint64()
        _check(lib().sqd_service_submit(self._h, program._h, first, last, ctypes.byref(j)))
        return int(j.value)

    def submit_many(self, program, ranges):
        """Jobs over the (first, last) ranges, queued together; their ids."""
        n = len(ranges)
        first = (ctypes.c_uint32 * max(n, 1))(*[a for a, _ in ranges])
        last = (ctypes.c_uint32 * max(n, 1))(*[b for _, b in ranges])
        ids = (ctypes.c_uint64 * max(n, 1))()
        _check(lib().sqd_service_submit_many(self._h, program._h, first, last, n, ids))
        return [int(ids[i]) for i in range(n)]

    def poll(self, job):
        done = ctypes.c_int()
        r = JobResult()
        _check(lib().sqd_service_poll(self._h, job, ctypes.byref(done), ctypes.byref(r)))
        return r.as_dict() if done.value else None

    def wait(self, job):
        r = JobResult()
        _check(lib().sqd_service_wait(self._h, job, ctypes.byref(r)))
        return r.as_dict()

    def stats(self):
        """Frontier bytes in flight and their peak, running rows and their peak,
        cores created and destroyed, jobs completed (sqd_service_get_stats)."""
        st = ServiceStats()
        _check(lib().sqd_service_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h:
            lib().sqd_service_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
