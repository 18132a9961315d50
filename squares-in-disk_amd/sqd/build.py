"""Build of the prover-core library (libsqd.so) and the CPU oracle.

Steps:
  1. compile the lowering tool (host C++, g++) and run it: one lowered program
     (gen/<key>.prog) and one generated HIP source (gen/<key>.hip) per lemma;
  2. compile the host sources with g++ and the HIP sources (runtime, arithmetic
     test kernel, generated programs) with hipcc for gfx950;
  3. link libsqd.so next to this file;
  4. build the CPU oracle (oracle/Makefile) for the tests and the CPU baseline.
"""
import hashlib
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent.parent          # squares-in-disk_amd/
ROOT = PKG.parent                                      # repository root
BUILD = PKG / "_build"
GEN = PKG / "gen"
LIB = PKG / "sqd" / "libsqd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", "g++")
ARCH = "gfx950"

INCLUDES = [ROOT / "include", PKG / "csrc", PKG / "host", PKG / "host" / "ivarp", PKG / "lemmas", GEN]
HOST_SOURCES = [PKG / "host/ivarp/expr.cpp", PKG / "host/ivarp/lower.cpp", PKG / "host/ivarp/program.cpp",
                PKG / "lemmas/lemmas.cpp"]
TOOL_SOURCES = HOST_SOURCES + [PKG / "host/ivarp/codegen.cpp", PKG / "host/tools/sqd_lower.cpp"]
HIP_SOURCES = [PKG / "csrc/runtime.hip", PKG / "csrc/service.hip", PKG / "csrc/arith_test.hip"]

CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall"]
HIPFLAGS = ["--offload-arch=" + ARCH, "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose=False, **kw):
    if verbose:
        print("+", " ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s%s" % (r.returncode, " ".join(str(c) for c in cmd),
                                                             r.stdout[-4000:], r.stderr[-8000:]))
    return r


def _inc():
    return ["-I" + str(p) for p in INCLUDES]


def _stamp(cmd, src):
    """Hash of a compile: the command, the source and every header it may include."""
    h = hashlib.sha256(" ".join(str(c) for c in cmd).encode())
    h.update(Path(src).read_bytes())
    for d in (PKG / "csrc", ROOT / "include", PKG / "host" / "ivarp", PKG / "lemmas"):
        for f in sorted(d.glob("*.h*")):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()


def build(verbose=False, jobs=None, incremental=False):
    """Compile everything; with incremental=True an object whose command, source
    and headers are unchanged since its last compile is reused."""
    jobs = jobs or max(1, min(16, os.cpu_count() or 1))
    BUILD.mkdir(exist_ok=True)
    GEN.mkdir(exist_ok=True)
    # 1. lowering tool and generated programs
    tool = BUILD / "sqd_lower"
    _run([CXX] + CXXFLAGS + _inc() + [str(s) for s in TOOL_SOURCES] + ["-o", str(tool)], verbose)
    for f in GEN.glob("*"):
        f.unlink()
    out = _run([tool, "gen", GEN], verbose).stdout
    gen_hip = sorted(GEN.glob("*.hip"))
    if verbose:
        print(out, end="")
    # 2. objects
    objs = []
    tasks = []
    def add(cmd, src, o):
        objs.append(o)
        st = o.with_suffix(".stamp")
        key = _stamp(cmd, src)
        if incremental and o.exists() and st.exists() and st.read_text() == key:
            return
        if st.exists():
            st.unlink()
        tasks.append((cmd, st, key))
    for s in HOST_SOURCES:
        o = BUILD / (s.stem + ".host.o")
        add([CXX] + CXXFLAGS + _inc() + ["-c", str(s), "-o", str(o)], s, o)
    for s in HIP_SOURCES + gen_hip:
        o = BUILD / (s.stem + ".hip.o")
        add([HIPCC] + HIPFLAGS + _inc() + ["-c", str(s), "-o", str(o)], s, o)

    def compile_one(task):
        cmd, st, key = task
        _run(cmd, verbose)
        st.write_text(key)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(compile_one, t) for t in tasks]:
            f.result()
    # 3. link
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs], verbose)
    os.replace(tmp, LIB)
    # 4. oracle (test infrastructure)
    _run(["make", "-C", str(ROOT / "oracle"), "-j%d" % min(jobs, 8)], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv)
    print("built", LIB)
