#!/usr/bin/env python3
"""Throughput of the GPU prover core on one MI355X.

Workload: the One Subcontainer Lemma (src/one_subcontainer.cpp in the
reference) proven in plain binary64 interval arithmetic with the reference
split policy: 128 initial dynamic cuboids on s_1, a 256 x 256 static grid on
(h_1, s_n), 8 generations -- 2,869,888 cuboids and 2,858,678 leaves per proof.

A step is one whole proof (initial box, initial queue, static levels, hull
reduction, final pass, dynamic splits) on the device; the program and its
constants are resident in HBM before timing starts.  Every step's counts are
checked against the published ones.  The line printed carries

  value       cuboids/s over the timed steps (wall clock around the steps)
  latency_ms  p50/p99 of the per-proof wall time
  roofline    the dominant kernel (largest summed device time), its
              algorithmic binary64 operations per launch (op-counted
              restatement of the same algorithm, per kernel kind) over its
              mean launch time from HIP events on the core's stream, against
              the MI355X FP64 vector peak; with the same mean launch time
              against the least time its VALU wave instructions (a
              SQ_INSTS_VALU count per launch from the counter summary that
              also gives the HBM traffic) need to issue, the bound that
              limits this interval kernel
  cpu_baseline the CPU restatement (oracle) timed on this host's cores on the
              same proof
  fp32_first_batched
              the same proofs through the proof service: the program's settings
              autotuned on batches (FP32 prefilter levels and children per lane,
              cuboids per round, blocks per launch), `--workers` workers each
              proving up to `--batch-jobs` queued proofs in the same launches
              (one grid row per proof) on its own stream; its
              cuboids/s, p50/p99 job latency and ratio to `value`, the fraction
              of static-level children the FP32 pass decided, and the
              FP32-versus-FP64 width ratio of each constraint on sampled boxes
  ratio_decomposition
              the batched leg's ratio split into its causes, on One and on
              Four, every configuration through the same service shape
              (`--workers` x `--batch-jobs`): plain binary64 in batches (no
              group test, no prefilter), the binary64 group test in batches,
              and the FP32-first service (autotuned); ratios to the plain
              batched rate
  legs        the other configurations of the metric, each with its counts
              checked against the published ones, cuboids/s, p50/p99, a
              roofline of its dominant kernel (binary64 operations per child
              from the op-counting restatement on the CPU sample below, times
              the children per launch the device counted, over the mean launch
              time) and a CPU baseline on a stated sample (a contiguous range
              of the first variable's initial subdivisions whose counts equal
              the device's on the same range):
                four_fp32_first   Four Subcontainer Lemma (37,245,328 cuboids),
                                  plain binary64 proofs and the FP32-first
                                  service with autotuned settings, the FP32
                                  decided fraction and width ratios
                more1_fp32_first  >=5 Subcontainers, s_n <= sigma, y_3 <= 0
                                  (1,056,154,896 cuboids) with HIP-graph
                                  rounds and the frontier in HBM, plain
                                  binary64 and FP32-first (autotuned)
  job_stream  proof jobs over all 11 statements (statement uniform, a
              contiguous 1/2^k slice of the first variable's initial
              subdivisions, k uniform in 0..6, seeded) through a service with a
              64 GiB device-memory budget, first run as one batch to
              measure capacity, then submitted with Poisson arrivals at 70 % of
              it: p50/p99 latency and cuboids/s; every job's counts are checked
              against the sum of its subtrees' (a sequential binary64 run)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--lemma one] [--workers 4] [--batch-jobs 16]
"""
import argparse
import json
import random
import os
import statistics
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "squares-in-disk_amd"))

METRIC = "cuboids/sec + p50/p99 proof-job latency, FP32-first batched vs plain FP64, 1 GPU"
FP64_PEAK_TFLOPS = 78.6          # MI355X binary64 vector peak
SIMDS, CLOCK_HZ, VALU_CYCLES = 1024, 2.4e9, 4   # MI355X: 256 CUs x 4 SIMDs; a wave64 VALU op holds a SIMD 4 cycles
PUBLISHED = {                    # (cuboids, leaves) from the reference's prover output
    "top1": (128, 128), "top2": (168, 168), "one": (2869888, 2858678), "two": (7181696, 7069587),
    "three": (3549856, 3440786), "four": (37245328, 35026303), "five": (3392, 3188), "six": (14848, 13928),
    "seven": (1808, 1703), "more1": (1056154896, 1022752013), "more2": (123127968, 115680431),
}
# published wall time of the reference's multithreaded CPU prover (unstated CPU, about 8 threads),
# prover_output_23_11_2020.txt:8-34
PUBLISHED_SECONDS = {"one": 1.13986, "four": 34.0994, "more1": 4068.32}
FP32_PEAK_TFLOPS = 157.3         # MI355X binary32 vector peak
WORKLOADS = {
    "one": "one_subcontainer_plain_fp64",
    "two": "two_subcontainer_plain_fp64",
    "three": "three_subcontainer_plain_fp64",
    "four": "four_subcontainer_plain_fp64",
    "more1": "more1_plain_fp64",
}
MODES = {0: "NONFINAL", 1: "FINAL", 2: "SPLIT", 3: "INIT"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--lemma", default="one", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-runs", type=int, default=3, help="timed oracle runs for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-settings", default=None, metavar="K=V,...",
                    help="fixed settings of the batched leg (e.g. fp32_prefilter=2,prefilter_group=8) instead of "
                         "autotuning them; for reproducible profiles")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "traffic_static_level_one_subcontainer.txt"),
                    help="HBM traffic summary (FETCH_SIZE/WRITE_SIZE per dispatch) to attach to the roofline")
    ap.add_argument("--workers", type=int, default=4, help="batches in flight in the batched legs")
    ap.add_argument("--batch-jobs", type=int, default=32, help="proofs per batch (one grid row each)")
    ap.add_argument("--stream-batch-jobs", type=int, default=4,
                    help="jobs per batch in the job-stream leg (11 programs share the memory budget)")
    ap.add_argument("--batched-proofs", type=int, default=4096)
    ap.add_argument("--stream-jobs", type=int, default=48)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the Four and More1 legs")
    ap.add_argument("--legs", default="four,more1", help="which of the Four and More1 legs to run")
    ap.add_argument("--no-graphs", action="store_true",
                    help="cores launch every round's kernels directly instead of replaying HIP graphs "
                         "(the same kernels; for counter collection)")
    ap.add_argument("--no-decomposition", action="store_true", help="skip the batched-ratio decomposition")
    ap.add_argument("--leg-proofs", type=int, default=64, help="batched Four proofs per configuration")
    ap.add_argument("--cpu-sample-cuboids", type=float, default=6e6,
                    help="about this many cuboids in the CPU sample of the Four and More1 legs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, one device-to-device copy of MIB MiB: a kernel of known traffic that "
                         "scales the profiler's FETCH_SIZE/WRITE_SIZE readings")
    return ap.parse_args()


def copy_calibration(mib):
    """One hipMemcpy device-to-device of `mib` MiB (a plain copy kernel)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(n)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise SystemExit("hipMemset failed")
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0 or hip.hipDeviceSynchronize() != 0:  # 3: device to device
            raise SystemExit("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def percentile(xs, q):
    s = sorted(xs)
    if not s:
        return None
    k = (len(s) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(s) - 1)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def cpu_baseline(program_text, runs):
    """Op counts per kernel kind and the CPU restatement's throughput on this host."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    threads = max(1, min(16, os.cpu_count() or 1))
    with tempfile.NamedTemporaryFile("w", suffix=".prog", delete=False) as f:
        f.write(program_text)
        path = f.name
    try:
        counted = pyoracle.run_cli(path, threads=threads, flops=True, timeout=600)
        timed = [pyoracle.run_cli(path, threads=threads, timeout=600) for _ in range(runs)]
    finally:
        os.unlink(path)
    secs = statistics.median(r["seconds"] for r in timed)
    return counted, {
        "value": timed[0]["num_cuboids"] / secs,
        "unit": "cuboids/s",
        "cores": threads,
        "kind": "port",
        "sample": "the whole proof (%d cuboids), median of %d runs, %d threads, %.3f s each"
                  % (timed[0]["num_cuboids"], runs, threads, secs),
    }


def fp32_op_counts(program_text):
    """Binary64 and binary32 operation counts per phase of the FP32-prefiltered proof (oracle)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    threads = max(1, min(16, os.cpu_count() or 1))
    with tempfile.NamedTemporaryFile("w", suffix=".prog", delete=False) as f:
        f.write(program_text)
        path = f.name
    try:
        return pyoracle.run_cli(path, threads=threads, flops=True, fp32=True, timeout=600)
    finally:
        os.unlink(path)


def read_counters(path, kernel_name):
    """The key = value lines of a counter summary of `kernel_name` (None if it covers another kernel)."""
    if not path or not Path(path).exists():
        return None
    vals = {}
    for line in Path(path).read_text().splitlines():
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    return vals if vals.get("kernel") == kernel_name else None


def read_traffic(path, kernel_name):
    """Bytes per launch of `kernel_name` from a traffic summary."""
    vals = read_counters(path, kernel_name)
    return float(vals["bytes_per_launch"]) if vals and "bytes_per_launch" in vals else None


def valu_issue_bound_ms(path, kernel_name):
    """The least time the launch's VALU wave instructions (SQ_INSTS_VALU) need at one
    issue per SIMD every 4 cycles on all 1024 SIMDs at 2.4 GHz."""
    vals = read_counters(path, kernel_name)
    if not vals or "valu_wave_instructions_per_launch" not in vals:
        return None
    return float(vals["valu_wave_instructions_per_launch"]) * VALU_CYCLES / (SIMDS * CLOCK_HZ) * 1e3


LEMMAS = ["top1", "top2", "one", "two", "three", "four", "five", "six", "seven", "more1", "more2"]


def initial_subdivisions(program):
    """The first dynamic variable's initial subdivisions (from the program text)."""
    for line in program.text().splitlines():
        f = line.split()
        if f and f[0] == "split" and f[4] == "1":
            return int(f[2])
    return 1


def fp32_first_batched(sqd, prog, expected, workers, batch_jobs, proofs, plain_value, counted32, fixed=None):
    """The batched FP32-first leg on `prog` (full proofs through the service)."""
    import ctypes
    svc_opts = sqd.ServiceOptions()
    sqd.lib().sqd_service_options_default(ctypes.byref(svc_opts))
    base = sqd.default_settings(jobs=batch_jobs, level_bytes=svc_opts.level_bytes,
                                frontier_capacity=svc_opts.frontier_capacity)
    if fixed:
        best = base
        for kv in fixed.split(","):
            k, v = kv.split("=")
            setattr(best, k.strip(), int(v, 0))
        tune = {"settings": {f: getattr(best, f) for f, _ in sqd.Settings._fields_}, "first": None, "last": None}
    else:
        best, tune = sqd.autotune(prog, base=base, fp32_first=True)   # on batches, as the service runs them
    tuned = {k: tune["settings"][k] for k in ("fp32_prefilter", "group_levels", "prefilter_group", "dequeue_buffer_size",
                                              "grid_blocks")}
    n = initial_subdivisions(prog)
    # decided fraction and the dominant prefiltered kernel's time on one core
    one = sqd.Settings.from_buffer_copy(best)
    one.jobs = 1
    with sqd.Core(prog, settings=one) as core:
        core.prove()
        core.set_profiling(True)
        for _ in range(20):
            core.prove()
        ks = core.kernel_stats()
        checked, decided = core.fp32_stats()
        core.set_profiling(False)
    with sqd.Service(workers=workers, batch_jobs=batch_jobs) as svc:
        svc.set_settings(prog, best)                          # the settings reported above
        warm_service(svc, prog, n, workers, batch_jobs)
        # all proofs handed over in one call, so that the timed region measures
        # the service, not the rate of one host thread's submissions
        t0 = time.perf_counter()
        ids = svc.submit_many(prog, [(0, n)] * proofs)
        res = [svc.wait(i) for i in ids]
        elapsed = time.perf_counter() - t0
    for r in res:
        c = r["counts"]
        if (c["num_cuboids"], c["num_leaf_cuboids"]) != expected or r["n_critical"]:
            raise SystemExit("count mismatch in the batched leg: %s" % c)
    lat = [r["latency_seconds"] for r in res]
    value = sum(r["counts"]["num_cuboids"] for r in res) / elapsed
    out = {
        "value": value, "unit": "cuboids/s", "ratio_vs_plain_fp64": value / plain_value,
        "workers": workers, "batch_jobs": batch_jobs, "proofs": proofs, "seconds": elapsed,
        "latency_ms": {"p50": percentile(lat, 0.5) * 1e3, "p99": percentile(lat, 0.99) * 1e3},
        "run_ms": {"p50": percentile([r["run_seconds"] for r in res], 0.5) * 1e3},
        "autotuned" if not fixed else "settings": tuned, "autotune_slice": [tune["first"], tune["last"]],
        "fp32_checked": checked, "fp32_decided": decided,
        "fp32_decided_fraction": decided / checked if checked else None,
        "single_core_kernel_ms_per_proof": {s["name"]: s["milliseconds"] / 20 for s in
                                            sorted(ks, key=lambda s: -s["milliseconds"])[:4]},
    }
    if counted32 is not None:
        out["fp32_ops_per_proof"] = {k: v for k, v in counted32.get("phases", {}).items()}
    return out


def warm_service(svc, prog, n, workers, batch_jobs):
    """Bursts of whole proofs until the service stops making cores (at least two
    bursts, at most eight): a burst deep enough that every worker starts a batch
    (a program gets a further batch while whole batches of its jobs wait), so
    that no core is made, allocated and captured inside the timed region."""
    made = -1
    for k in range(8):
        for i in svc.submit_many(prog, [(0, n)] * (4 * workers * batch_jobs)):
            svc.wait(i)
        now = svc.stats()["cores_created"]
        if k >= 1 and now == made:
            break
        made = now


def service_batches(sqd, prog, settings, expected, workers, batch_jobs, proofs):
    """`proofs` whole proofs of `prog` through a service whose cores use `settings`
    (`workers` workers, `batch_jobs` proofs per batch); every proof's counts are
    checked.  Cuboids/s over the submitted proofs and p50/p99 job latency."""
    n = initial_subdivisions(prog)
    with sqd.Service(workers=workers, batch_jobs=batch_jobs) as svc:
        svc.set_settings(prog, settings)
        warm_service(svc, prog, n, workers, batch_jobs)
        # all proofs handed over in one call, so that the timed region measures
        # the service, not the rate of one host thread's submissions
        t0 = time.perf_counter()
        ids = svc.submit_many(prog, [(0, n)] * proofs)
        res = [svc.wait(i) for i in ids]
        elapsed = time.perf_counter() - t0
    for r in res:
        c = r["counts"]
        if (c["num_cuboids"], c["num_leaf_cuboids"]) != expected or r["n_critical"]:
            raise SystemExit("count mismatch in a batched run: %s" % c)
    lat = [r["latency_seconds"] for r in res]
    return {"value": sum(r["counts"]["num_cuboids"] for r in res) / elapsed, "unit": "cuboids/s",
            "proofs": proofs, "seconds": elapsed,
            "latency_ms": {"p50": percentile(lat, 0.5) * 1e3, "p99": percentile(lat, 0.99) * 1e3}}


def service_settings(sqd, batch_jobs, **kw):
    """Settings of a service core (the service's buffer sizes) with the given fields."""
    import ctypes
    o = sqd.ServiceOptions()
    sqd.lib().sqd_service_options_default(ctypes.byref(o))
    s = sqd.default_settings(jobs=batch_jobs, level_bytes=o.level_bytes, frontier_capacity=o.frontier_capacity)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def decompose(sqd, key, workers, batch_jobs, proofs, fp32_first=None, fp32_settings=None):
    """Plain binary64, binary64 group test and FP32-first, all in batches of the
    same service shape, on whole proofs of `key`; fp32_first: an already measured
    FP32-first result (with its settings) to reuse, else tuned and run here."""
    prog = sqd.Program.from_lemma(key)
    expected = PUBLISHED[key]
    plain = service_batches(sqd, prog, service_settings(sqd, batch_jobs, fp32_prefilter=0, group_levels=0), expected,
                            workers, batch_jobs, proofs)
    grouped = service_batches(sqd, prog, service_settings(sqd, batch_jobs, fp32_prefilter=0, group_levels=0xFFFFFFFF,
                                                          prefilter_group=16), expected, workers, batch_jobs, proofs)
    if fp32_first is None:
        if fp32_settings is None:
            fp32_settings, _ = sqd.autotune(prog, base=service_settings(sqd, batch_jobs), fp32_first=True)
        fp32_first = service_batches(sqd, prog, fp32_settings, expected, workers, batch_jobs, proofs)
        fp32_first["settings"] = {k: getattr(fp32_settings, k) for k in
                                  ("fp32_prefilter", "group_levels", "prefilter_group", "dequeue_buffer_size",
                                   "grid_blocks")}
    return {
        "lemma": key, "workers": workers, "batch_jobs": batch_jobs, "proofs": proofs,
        "plain_batched": plain, "grouped_binary64_batched": dict(grouped, settings={"group_levels": "all",
                                                                                     "prefilter_group": 16}),
        "fp32_first_batched": fp32_first,
        "ratio_grouped_vs_plain_batched": grouped["value"] / plain["value"],
        "ratio_vs_plain_batched": fp32_first["value"] / plain["value"],
        "ratio_fp32_first_vs_grouped": fp32_first["value"] / grouped["value"],
    }


def profiled_proof(sqd, core, expected):
    """One profiled proof: the kernel statistics (time, launches, children) of every kind."""
    core.set_profiling(True)
    r = core.prove()
    stats = core.kernel_stats()
    core.set_profiling(False)
    c = r["counts"]
    if (c["num_cuboids"], c["num_leaf_cuboids"]) != expected or r["n_critical"]:
        raise SystemExit("count mismatch in a profiled proof: %s" % c)
    return stats


def dominant(stats):
    return max((s for s in stats if s["mode"] >= 0), key=lambda s: s["milliseconds"])


def timed_proofs(core, expected, steps):
    lat, total = [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        r = core.prove()
        lat.append(time.perf_counter() - s0)
        c = r["counts"]
        if (c["num_cuboids"], c["num_leaf_cuboids"]) != expected or r["n_critical"]:
            raise SystemExit("count mismatch: %s" % c)
        total += c["num_cuboids"]
    elapsed = time.perf_counter() - t0
    return {"value": total / elapsed, "unit": "cuboids/s", "proofs": steps, "seconds": elapsed,
            "latency_ms": {"p50": percentile(lat, 0.5) * 1e3, "p99": percentile(lat, 0.99) * 1e3}}


def sample_range(sqd, prog, target):
    """The contiguous range of the first variable's initial subdivisions whose
    proof has the cuboid count closest to `target` (in ratio; per-subdivision
    counts from the device; ties go to the first range)."""
    import math
    n = initial_subdivisions(prog)
    with sqd.Core(prog) as core:
        per = [core.prove_range(i, i + 1)["counts"]["num_cuboids"] for i in range(n)]
    pre = [0]
    for v in per:
        pre.append(pre[-1] + v)
    best = None
    for a in range(n):
        for b in range(a + 1, n + 1):
            t = pre[b] - pre[a]
            if t <= 0:
                continue
            d = abs(math.log(t / target))
            if best is None or d < best[0]:
                best = (d, a, b, t)
    return best[1], best[2], best[3]


def oracle_sample(sqd, prog, rng, fp32=False):
    """The CPU restatement on the initial range `rng` (None: the whole proof): a
    timed run and an op-counting run (binary32 too with fp32)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    threads = max(1, min(16, os.cpu_count() or 1))
    with tempfile.NamedTemporaryFile("w", suffix=".prog", delete=False) as f:
        f.write(prog.text())
        path = f.name
    try:
        timed = pyoracle.run_cli(path, threads=threads, init_range=rng, timeout=900)
        counted = pyoracle.run_cli(path, threads=threads, init_range=rng, flops=True, timeout=900)
        counted32 = pyoracle.run_cli(path, threads=threads, init_range=rng, flops=True, fp32=True,
                                     timeout=900) if fp32 else None
    finally:
        os.unlink(path)
    return threads, timed, counted, counted32


def leg_roofline(stat, phases, fp32=False):
    """Roofline of a level kernel kind: operations per child from the op-counted
    restatement (phase "mode.level": [binary64 ops, children, binary32 ops]) times
    the children per launch the device counted, over the mean launch time."""
    ph = phases.get("%d.%d" % (stat["mode"], stat["level"]))
    if not ph or not ph[1] or not stat["launches"] or not stat["children"]:
        return None
    mean_ms = stat["milliseconds"] / stat["launches"]
    children = stat["children"] / stat["launches"]
    f64 = ph[0] / ph[1] * children
    f32 = ph[2] / ph[1] * children if fp32 else 0.0
    t = mean_ms * 1e-3
    a64, a32 = f64 / t / 1e12, f32 / t / 1e12
    frac = a64 / FP64_PEAK_TFLOPS + a32 / FP32_PEAK_TFLOPS       # share of the launch the ops need at peak
    peak = (a64 + a32) / frac if frac else FP64_PEAK_TFLOPS
    return {"bound": "valu", "achieved": a64 + a32, "peak": peak, "unit": "TFLOP/s", "frac": frac,
            "traffic": None, "kernel": stat["name"], "mean_launch_ms": mean_ms,
            "children_per_launch": children, "flops_per_launch": f64 + f32,
            "binary64_ops_per_child": ph[0] / ph[1], "binary32_ops_per_child": ph[2] / ph[1] if fp32 else 0.0,
            "achieved_binary64": a64, "achieved_binary32": a32}


def lemma_leg(sqd, key, plain_steps, sample_cuboids, workers, batch_jobs, proofs, cpu=True, batched=True):
    """One configuration of the metric beyond One (see the module docstring)."""
    prog = sqd.Program.from_lemma(key)
    expected = PUBLISHED[key]
    out = {"lemma": key, "cuboids_per_proof": expected[0], "leaves_per_proof": expected[1],
           "reference_published_seconds": PUBLISHED_SECONDS.get(key)}
    # plain binary64, reference split policy, HIP-graph rounds, frontier in HBM
    with sqd.Core(prog, settings=sqd.default_settings(generation_count=8)) as core:
        core.prove()
        stats = profiled_proof(sqd, core, expected)
        core.prove()                                          # captures the round graphs
        out["plain_fp64"] = timed_proofs(core, expected, plain_steps)
    dom = dominant(stats)
    out["plain_fp64"]["kernel_ms_per_proof"] = {s["name"]: s["milliseconds"] for s in
                                                sorted(stats, key=lambda s: -s["milliseconds"])[:4]}
    # FP32-first on a core of one job (autotuned), HIP-graph rounds
    best, tune = sqd.autotune(prog, base=sqd.default_settings(), fp32_first=True)
    with sqd.Core(prog, settings=best) as core:
        core.prove()
        stats32 = profiled_proof(sqd, core, expected)
        checked, decided = core.fp32_stats()
        core.prove()
        out["fp32_first"] = timed_proofs(core, expected, plain_steps)
    dom32 = dominant(stats32)
    out["fp32_first"].update({
        "settings": {k: tune["settings"][k] for k in ("fp32_prefilter", "group_levels", "prefilter_group",
                                                      "dequeue_buffer_size", "grid_blocks")},
        "fp32_checked": checked, "fp32_decided": decided,
        "fp32_decided_fraction": decided / checked if checked else None,
        "kernel_ms_per_proof": {s["name"]: s["milliseconds"] for s in
                                sorted(stats32, key=lambda s: -s["milliseconds"])[:4]},
        "ratio_vs_plain_fp64": None})
    out["fp32_first"]["ratio_vs_plain_fp64"] = out["fp32_first"]["value"] / out["plain_fp64"]["value"]
    out["value"] = out["fp32_first"]["value"]
    if batched:
        out["batched"] = decompose(sqd, key, workers, batch_jobs, proofs)
        out["value"] = max(out["value"], out["batched"]["fp32_first_batched"]["value"])
    out["width_ratio_per_constraint"] = width_sample(sqd, prog, n_boxes=4096)
    if cpu:
        a, b, gpu_cuboids = sample_range(sqd, prog, sample_cuboids)
        with sqd.Core(prog) as core:
            g = core.prove_range(a, b)["counts"]
        threads, timed, counted, counted32 = oracle_sample(sqd, prog, (a, b), fp32=True)
        if (timed["num_cuboids"], timed["num_leaf_cuboids"]) != (g["num_cuboids"], g["num_leaf_cuboids"]):
            raise SystemExit("%s: oracle sample differs from the device on [%d, %d)" % (key, a, b))
        out["cpu_baseline"] = {
            "value": timed["num_cuboids"] / timed["seconds"], "unit": "cuboids/s", "cores": threads, "kind": "port",
            "sample": "initial subdivisions [%d, %d) of the first variable: %d cuboids (equal to the device's), "
                      "%.2f s with %d threads" % (a, b, timed["num_cuboids"], timed["seconds"], threads)}
        out["roofline"] = attach_counters(leg_roofline(dom, counted["phases"]), key)
        out["roofline_fp32_first"] = attach_counters(leg_roofline(dom32, counted32["phases"],
                                                                  fp32=bool(best.fp32_prefilter)), key)
        out["op_count_sample"] = "the same range, op-counting restatement"
    return out


def attach_counters(roof, key):
    """HBM traffic and the VALU-issue bound of a leg's dominant kernel from its
    counter summary under profiles/ (traffic_<kernel kind>_<statement>.txt), when
    one covers that kernel."""
    if not roof:
        return roof
    kind = roof["kernel"].split("<")[0]
    path = ROOT / "profiles" / ("traffic_%s_%s.txt" % (kind, key))
    traffic = read_traffic(path, roof["kernel"])
    issue_ms = valu_issue_bound_ms(path, roof["kernel"])
    if traffic is not None:
        roof["traffic"] = traffic
    if issue_ms is not None:
        roof["valu_issue_bound_ms"] = issue_ms
        roof["valu_issue_frac"] = issue_ms / roof["mean_launch_ms"]
    return roof


def width_sample(sqd, prog, n_boxes=8192, seed=3):
    """FP32-versus-FP64 width ratio per constraint on boxes of the static grid."""
    import numpy as np
    rng = np.random.default_rng(seed)
    with sqd.Core(prog) as core:
        qlb, qub, _ = core.initial_queue()
        pick = rng.integers(0, qlb.shape[1], n_boxes)
        lb, ub = qlb[:, pick].copy(), qub[:, pick].copy()
        for line in prog.text().splitlines():                 # one random cell of every static variable
            f = line.split()
            if f and f[0] == "split" and f[4] == "0":
                a, subs = int(f[1]), int(f[3])
                cell = rng.integers(0, subs, n_boxes)
                w = (ub[a] - lb[a]) / subs
                lb[a], ub[a] = lb[a] + cell * w, lb[a] + (cell + 1) * w
        st = core.width_stats(lb, ub)
    return {"c%d" % s["constraint"]: {"mean_ratio": round(s["mean_ratio"], 4), "max_ratio": round(s["max_ratio"], 4),
                                       "samples": s["samples"]} for s in st if s["samples"]}


def job_stream(sqd, jobs_n, workers, batch_jobs, seed):
    """Poisson stream of proof jobs over all statements (see the module docstring)."""
    rng = random.Random(seed)
    progs = {k: sqd.Program.from_lemma(k) for k in LEMMAS}
    init = {k: initial_subdivisions(p) for k, p in progs.items()}
    jobs = []
    for _ in range(jobs_n):
        k = rng.choice(LEMMAS)
        e = rng.randint(0, 6)
        size = init[k] >> e
        first = rng.randrange(0, init[k] // size) * size
        jobs.append((k, first, first + size))
    # reference counts: each job alone, one at a time, binary64
    alone = {}
    with sqd.Service(workers=1, fp32_prefilter=0) as svc:
        for j in set(jobs):
            alone[j] = svc.wait(svc.submit(progs[j[0]], j[1], j[2]))["counts"]
    with sqd.Service(workers=workers, autotune=1, batch_jobs=batch_jobs, memory_budget=64 << 30) as svc:
        for k in LEMMAS:                                      # tune and create a core per statement
            svc.wait(svc.submit(progs[k], 0, 1))
        t0 = time.perf_counter()
        ids = [svc.submit(progs[k], a, b) for k, a, b in jobs]
        batch = [svc.wait(i) for i in ids]
        makespan = time.perf_counter() - t0
        rate = 0.7 * len(jobs) / makespan                     # 70 % of the measured capacity
        arrivals, t = [], 0.0
        for _ in jobs:
            t += rng.expovariate(rate)
            arrivals.append(t)
        t0 = time.perf_counter()
        ids = []
        for (k, a, b), at in zip(jobs, arrivals):
            d = at - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids.append(svc.submit(progs[k], a, b))
        res = [svc.wait(i) for i in ids]
        elapsed = time.perf_counter() - t0
    for j, r, rb in zip(jobs, res, batch):
        if r["counts"] != alone[j] or rb["counts"] != alone[j] or r["n_critical"]:
            raise SystemExit("job stream: %s differs from its sequential run" % (j,))
    lat = [r["latency_seconds"] for r in res]
    cub = sum(r["counts"]["num_cuboids"] for r in res)
    return {
        "jobs": len(jobs), "seed": seed, "workers": workers, "batch_jobs": batch_jobs, "memory_budget_gib": 64,
        "arrival_rate_per_s": rate,
        "capacity_jobs_per_s": len(jobs) / makespan,
        "latency_ms": {"p50": percentile(lat, 0.5) * 1e3, "p99": percentile(lat, 0.99) * 1e3},
        "cuboids_per_s": cub / elapsed, "cuboids": cub, "seconds": elapsed,
        "same_as_alone": True,
    }


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py: this prover core runs on one GPU; multi-GPU runs (--gpus %d, WORLD_SIZE %d) are not "
              "part of this build" % (a.gpus, world), file=sys.stderr)
        sys.exit(2)
    import sqd
    if a.no_graphs:                      # every core this run makes, through the binding's defaults
        plain_defaults = sqd.default_settings

        def no_graph_defaults(**kw):
            st = plain_defaults(**kw)
            st.graphs = 0
            return st
        sqd.default_settings = no_graph_defaults

    prog = sqd.Program.from_lemma(a.lemma)
    expected = PUBLISHED[a.lemma]
    core = sqd.Core(prog, settings=sqd.default_settings(generation_count=8))
    try:
        # warmup, timing every launch once to find the dominant kernel kind
        core.set_profiling(True)
        for _ in range(max(1, a.warmup)):
            r = core.prove()
            c = r["counts"]
            if (c["num_cuboids"], c["num_leaf_cuboids"]) != expected or r["n_critical"]:
                raise SystemExit("count mismatch in warmup: %s" % c)
        warm = core.kernel_stats()
        dom = max((s for s in warm if s["mode"] >= 0), key=lambda s: s["milliseconds"])
        # timed steps: only the dominant kernel is bracketed by events
        core.set_profiling(True, only=(dom["mode"], dom["level"]))
        lat = []
        t0 = time.perf_counter()
        total_cuboids = 0
        for _ in range(a.steps):
            s0 = time.perf_counter()
            r = core.prove()
            lat.append(time.perf_counter() - s0)
            c = r["counts"]
            if (c["num_cuboids"], c["num_leaf_cuboids"]) != expected or r["n_critical"]:
                raise SystemExit("count mismatch: %s" % c)
            total_cuboids += c["num_cuboids"]
        elapsed = time.perf_counter() - t0
        stats = core.kernel_stats()
        core.set_profiling(False)
    finally:
        core.close()

    counted, cpu, counted32 = (None, None, None)
    if not a.no_cpu_baseline:
        counted, cpu = cpu_baseline(prog.text(), a.cpu_runs)
        counted32 = fp32_op_counts(prog.text())
    value = total_cuboids / elapsed
    batched = None if a.no_batched else fp32_first_batched(sqd, prog, expected, a.workers, a.batch_jobs,
                                                           a.batched_proofs, value, counted32, a.batched_settings)
    if batched is not None:
        batched["width_ratio_per_constraint"] = width_sample(sqd, prog)
    decomposition = {}
    if batched is not None and not a.no_decomposition:
        mine = {k: batched[k] for k in ("value", "unit", "proofs", "seconds", "latency_ms")}
        mine["settings"] = batched.get("autotuned", batched.get("settings"))
        decomposition[a.lemma] = decompose(sqd, a.lemma, a.workers, a.batch_jobs, a.batched_proofs, fp32_first=mine)
        batched["ratio_vs_plain_batched"] = decomposition[a.lemma]["ratio_vs_plain_batched"]
    legs = {}
    which = set() if a.no_legs else set(a.legs.split(","))
    if "four" in which:
        legs["four_fp32_first"] = lemma_leg(sqd, "four", 10, a.cpu_sample_cuboids, a.workers, a.batch_jobs,
                                            a.leg_proofs, cpu=not a.no_cpu_baseline,
                                            batched=not a.no_decomposition)
        if "batched" in legs["four_fp32_first"]:
            decomposition["four"] = legs["four_fp32_first"]["batched"]
    if "more1" in which:
        legs["more1_fp32_first"] = lemma_leg(sqd, "more1", 3, a.cpu_sample_cuboids, a.workers, a.batch_jobs,
                                             0, cpu=not a.no_cpu_baseline, batched=False)
    stream = None if a.no_stream else job_stream(sqd, a.stream_jobs, a.workers, a.stream_batch_jobs, a.seed)

    dom = next(s for s in stats if s["mode"] == dom["mode"] and s["level"] == dom["level"])
    mean_ms = dom["milliseconds"] / dom["launches"]
    roofline = None
    if counted is not None:
        ph = counted["phases"].get("%d.%d" % (dom["mode"], dom["level"]), [0, 0, 0])
        flops_per_proof, cuboids_per_proof = ph[0], ph[1]
        launches_per_proof = dom["launches"] / a.steps
        flops_per_launch = flops_per_proof / launches_per_proof
        achieved = flops_per_launch / (mean_ms * 1e-3) / 1e12
        traffic = read_traffic(a.traffic, dom["name"])
        issue_ms = valu_issue_bound_ms(a.traffic, dom["name"])
        roofline = {
            "bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "kernel": dom["name"], "mean_launch_ms": mean_ms, "launches_per_proof": launches_per_proof,
            "flops_per_launch": flops_per_launch,
            "cuboids_per_launch": cuboids_per_proof / launches_per_proof,
            "flops_per_cuboid": flops_per_proof / cuboids_per_proof if cuboids_per_proof else None,
            "valu_issue_bound_ms": issue_ms,
            "valu_issue_frac": issue_ms / mean_ms if issue_ms else None,
        }
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cuboids/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "lemma-defined (no input data)",
        "config": {"workload": WORKLOADS[a.lemma], "lemma": a.lemma, "arithmetic": "plain binary64 intervals",
                   "split_policy": "reference", "generations": 8, "cuboids_per_proof": expected[0],
                   "leaves_per_proof": expected[1]},
        "latency_ms": {"p50": percentile(lat, 0.5) * 1e3, "p99": percentile(lat, 0.99) * 1e3},
        "warmup_kernel_ms_per_step": {s["name"]: s["milliseconds"] / max(1, a.warmup) for s in
                                      sorted(warm, key=lambda s: -s["milliseconds"])},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32_first_batched": batched,
        "ratio_decomposition": decomposition or None,
        "legs": legs or None,
        "job_stream": stream,
    }
    print(json.dumps(line), flush=True)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)


if __name__ == "__main__":
    main()
