"""The reference-side binding (include/sqd_hip_prover_core.hpp) compiles and links.

The adapter derives from ivarp's BasicProverCore and uses ProofDriverQueueEntry,
CuboidCounts and ProverSettings (prover/basic_prover_core.hpp:29-67,
prover/proof_driver.hpp:34-78, prover/prover_settings.hpp).  This test supplies
minimal stand-ins of its own with the members the adapter touches, instantiates
the adapter for a three-argument statement, and links it against libsqd.so, so a
change of the C ABI that breaks the adapter fails here, without a GPU.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB_DIR = ROOT / "squares-in-disk_amd" / "sqd"

STANDINS = r"""
#include <cstddef>
#include <cstdio>
#include <vector>
namespace ivarp {
struct ProverSettings { int generation_count = 8; std::size_t dequeue_buffer_size = 0; };
struct CuboidCounts {
    std::size_t num_cuboids = 0, num_leaf_cuboids = 0, num_critical_cuboids = 0, num_repeated_nodes = 0;
};
template<typename NumberType, std::size_t NumArgs> struct ProofDriverQueueEntry {
    NumberType elements[NumArgs];
    std::size_t depth;
};
template<typename ProverInputType, typename OnCritical> class BasicProverCore {
public:
    virtual void set_settings(const ProverSettings* s) { settings = s; }
    void set_on_critical(const OnCritical* c) noexcept { on_critical = c; }
protected:
    BasicProverCore() noexcept = default;
    ~BasicProverCore() = default;
    const OnCritical* on_critical{nullptr};
    const ProverSettings* settings{nullptr};
};
}  // namespace ivarp
#include "sqd_hip_prover_core.hpp"

struct IDouble {
    double l = 0, u = 0;
    IDouble() = default;
    IDouble(double a, double b) : l(a), u(b) {}
    double lb() const { return l; }
    double ub() const { return u; }
};
struct Context { using NumberType = IDouble; };
struct Input { using Context = ::Context; static constexpr std::size_t num_args = 3; };
struct OnCritical {
    std::size_t* n;
    void operator()(const Context&, const IDouble*) const { ++*n; }
};
template class ivarp::HipProverCore<Input, OnCritical>;

int main(int argc, char**) {
    if(argc > 1) {   // never run by the test: the library is only linked
        ivarp::HipProverCore<Input, OnCritical> core("one");
        ivarp::ProverSettings s;
        std::size_t n = 0;
        OnCritical oc{&n};
        core.set_settings(&s);
        core.set_on_critical(&oc);
        core.initialize(1);
        core.initialize_per_thread(0, 1);
        std::vector<ivarp::ProofDriverQueueEntry<IDouble, 3>> in(1), out;
        core.handle_cuboids_nonfinal(0, in, &out);
        core.handle_cuboids_final(0, in);
        std::printf("%zu\n", n);
    }
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_adapter_compiles_and_links(tmp_path):
    if not (LIB_DIR / "libsqd.so").exists():
        pytest.skip("libsqd.so is not built")
    src = tmp_path / "adapter.cpp"
    src.write_text(STANDINS)
    exe = tmp_path / "adapter"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
           "-L", str(LIB_DIR), "-lsqd", "-Wl,--allow-shlib-undefined", "-Wl,-rpath," + str(LIB_DIR)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
