"""Every critical cuboid across the boundary (on_critical semantics).

The reference hands each critical to its on_critical callback
(cpu_prover/cpu_prover_core.hpp:137-149); its GPU core re-launches until the
output buffer holds them all (with_cuda/ivarp_cuda/gpu_prover_core.cuh:293-316).
Here a critical sink receives every critical of a proof or a final pass in
chunks, with a device list far smaller than the criticals; the boxes must be bit
for bit the oracle's (known-answer system of ivarp/test/expected_criticals2.cpp).
"""
import numpy as np
import pytest

import pyoracle
import sqd

pytestmark = pytest.mark.gpu


def _rows(lb, ub):
    """Boxes as rows of bit patterns (lb_0, ub_0, lb_1, ub_1, ...), sorted."""
    n_args = lb.shape[0]
    m = np.empty((lb.shape[1], 2 * n_args), dtype=np.uint64)
    m[:, 0::2] = lb.T.view(np.uint64) if lb.dtype == np.float64 else lb.T
    m[:, 1::2] = ub.T.view(np.uint64) if ub.dtype == np.float64 else ub.T
    return m[np.lexsort(m.T[::-1])]


def _oracle_rows(crit):
    rows = np.ascontiguousarray(crit.reshape(len(crit), -1)).view(np.uint64)
    return rows[np.lexsort(rows.T[::-1])]


class Collect:
    def __init__(self):
        self.lb, self.ub, self.calls = [], [], 0

    def __call__(self, lb, ub):
        self.lb.append(lb)
        self.ub.append(ub)
        self.calls += 1

    def boxes(self, n_args):
        if not self.lb:
            return np.zeros((n_args, 0)), np.zeros((n_args, 0))
        return np.hstack(self.lb), np.hstack(self.ub)


@pytest.mark.parametrize("graphs", [0, 1])
def test_sink_receives_every_critical_of_a_proof(program_text, graphs):
    # ec2 at 2 generations: 262,144 criticals, with a device list of 4,096
    prog = sqd.Program.from_lemma("ec2")
    o_counts, o_crit = pyoracle.Oracle(program_text("ec2")).run(generations=2, threads=8, max_criticals=1 << 19)
    assert o_counts["num_critical_cuboids"] == 262144
    s = sqd.default_settings(generation_count=2, critical_capacity=4096, graphs=graphs)
    with sqd.Core(prog, settings=s) as core:
        for rep in range(2):                       # the second proof replays graphs when they are on
            sink = Collect()
            core.set_critical_sink(sink)
            r = core.prove()
            assert r["counts"] == o_counts and r["n_critical"] == 262144
            lb, ub = sink.boxes(prog.nargs)
            assert lb.shape[1] == 262144 and sink.calls >= 262144 // 4096
            assert np.array_equal(_rows(lb, ub), _oracle_rows(o_crit))


def test_sink_with_prefilter_matches_oracle(program_text):
    # the same with the binary32 prefilter on every static level: the criticals
    # are the oracle's binary32-prefiltered ones
    prog = sqd.Program.from_lemma("ec2")
    o = pyoracle.Oracle(program_text("ec2"), fp32=True)
    o_counts, o_crit = o.run(generations=2, threads=8, max_criticals=1 << 19)
    s = sqd.default_settings(generation_count=2, critical_capacity=1000, fp32_prefilter=sqd.FP32_ALL_LEVELS)
    with sqd.Core(prog, settings=s) as core:
        sink = Collect()
        core.set_critical_sink(sink)
        r = core.prove()
    assert r["counts"] == o_counts
    lb, ub = sink.boxes(prog.nargs)
    assert np.array_equal(_rows(lb, ub), _oracle_rows(o_crit))


def test_without_sink_overflow_is_a_capacity_error():
    prog = sqd.Program.from_lemma("ec2")
    s = sqd.default_settings(generation_count=2, critical_capacity=4096)
    with sqd.Core(prog, settings=s) as core:
        with pytest.raises(sqd.SqdError) as e:
            core.prove()
        assert e.value.code == -4 and "critical_capacity" in str(e.value)
        # the core stays usable: the same proof with a sink receives every critical
        sink = Collect()
        core.set_critical_sink(sink)
        r = core.prove()
        assert r["n_critical"] == 262144 and sink.boxes(prog.nargs)[0].shape[1] == 262144


def test_final_pass_sink_in_parts(program_text):
    # handle_cuboids_final on 64 x 64 boxes of ec2 (ivarp/test/expected_criticals2.cpp:99-148):
    # 262,144 criticals through a device list of 512, in parts of 64 inputs of the last level
    prog = sqd.Program.from_lemma("ec2")
    k = 64
    pts = np.array([i / k for i in range(k + 1)])
    w = np.repeat(np.arange(k), k)
    x = np.tile(np.arange(k), k)
    n = k * k
    lb = np.vstack([pts[w], pts[x], np.zeros(n), np.zeros(n)])
    ub = np.vstack([pts[w + 1], pts[x + 1], np.ones(n), np.ones(n)])
    ol, ou, oc = pyoracle.Oracle(program_text("ec2")).final(lb, ub, cap=1 << 19)
    with sqd.Core(prog, settings=sqd.default_settings(critical_capacity=512)) as core:
        sink = Collect()
        core.set_critical_sink(sink)
        cl, cu, c = core.final(lb, ub, capacity=16)
        assert c == oc and c["num_critical_cuboids"] == n * 64
        assert cl.shape[1] == 0                      # nothing copied to the call's own arrays
        glb, gub = sink.boxes(prog.nargs)
        assert np.array_equal(_rows(glb, gub), _rows(ol, ou))
        core.set_critical_sink(None)                 # without the sink: more than the list holds fails
        with pytest.raises(sqd.SqdError):
            core.final(lb, ub, capacity=16)
