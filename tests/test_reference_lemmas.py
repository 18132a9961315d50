"""The reference's own lemma sources compile unchanged against this project's
ivarp surface (squares-in-disk_amd/compat: <ivarp/math_fn.hpp>,
<ivarp/prover.hpp>, <ivarp/run_prover.hpp>, <ivarp/critical_printer.hpp>,
<ivarp/progress_printer.hpp> over the run-time DSL and the C ABI) and lower to
exactly the programs the GPU core proves for the built-in statements.

Each file of the reference's src/ is included unchanged into a small
translation unit that prints the lowered program of its prover_input (the
files' inputs are namespace-scope constants of internal linkage); the whole
unit, including the file's run_* driver, links against libsqd.so.  Skipped
where the reference is absent (it never travels with this repository).
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "squares-in-disk_amd"
REF_SRC = Path("/root/reference/src")

# (source file, namespace of its proof, built-in key)
CASES = [
    ("top_packing.cpp", "top_packing::proof1", "top1"),
    ("top_packing.cpp", "top_packing::proof2", "top2"),
    ("one_subcontainer.cpp", "one_subcontainer_proof1", "one"),
    ("two_subcontainer.cpp", "two_subcontainers_proof1", "two"),
    ("three_subcontainer.cpp", "three_subcontainers_proof1", "three"),
    ("four_subcontainer.cpp", "four_subcontainers_proof1", "four"),
    ("five_subcontainer.cpp", "five_subcontainers_proof1", "five"),
    ("six_subcontainer.cpp", "six_subcontainers_proof1", "six"),
    ("seven_subcontainer.cpp", "seven_subcontainers_proof1", "seven"),
    ("more_subcontainers.cpp", "more_subcontainers_proof1", "more1"),
    ("more_subcontainers.cpp", "more_subcontainers_proof2", "more2"),
]

pytestmark = [pytest.mark.skipif(not REF_SRC.is_dir(), reason="the reference is not present"),
              pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")]


def _body(text):
    """The program text without its name line (the built-in statements carry a title)."""
    return [line for line in text.splitlines() if not line.startswith("name ")]


@pytest.fixture(scope="module")
def host_objects(tmp_path_factory):
    """The lowering's host sources, compiled once."""
    d = tmp_path_factory.mktemp("compat")
    objs = []
    for src in ("expr.cpp", "lower.cpp", "program.cpp"):
        o = d / (src + ".o")
        r = subprocess.run(["g++", "-std=c++17", "-O1", "-c", "-I", str(PKG / "host"), "-I", str(PKG / "host/ivarp"),
                            str(PKG / "host/ivarp" / src), "-o", str(o)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        objs.append(str(o))
    return d, objs


@pytest.mark.parametrize("src,ns,key", CASES, ids=[c[2] for c in CASES])
def test_reference_lemma_compiles_unchanged_and_lowers_identically(host_objects, src, ns, key):
    if not (PKG / "sqd" / "libsqd.so").exists():
        pytest.skip("libsqd.so is not built")
    import sqd
    d, objs = host_objects
    unit = d / ("unit_%s.cpp" % key)
    unit.write_text('#include "%s"\n#include <iostream>\n'
                    'int main() { %s::input.program.write(std::cout); return 0; }\n' % (REF_SRC / src, ns))
    exe = d / ("unit_%s" % key)
    cmd = ["g++", "-std=c++17", "-O1", "-I", str(PKG / "compat"), "-I", str(PKG / "host"), "-I", str(PKG / "host/ivarp"),
           "-I", str(ROOT / "include"), "-I", str(REF_SRC), str(unit)] + objs + \
          ["-o", str(exe), "-L", str(PKG / "sqd"), "-lsqd", "-Wl,--allow-shlib-undefined",
           "-Wl,-rpath," + str(PKG / "sqd")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    out = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec " + str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-2000:]
    assert _body(out.stdout) == _body(sqd.Program.from_lemma(key).text())
