"""Lemma-level enclosure on the device: the interval of every bound expression
and the IBool of every constraint of each statement, evaluated by its generated
code (compile-time sign shortcuts, lazy if_then_else, domain clamping) in
binary64 and, on boxes converted outward, in binary32 (the prefilter), against
the real values at sampled points of boxes the proofs visit (checks in
tests/enclosure_check.py); and the same values bit for bit the oracle's."""
import numpy as np
import pytest

import enclosure_check as ec
import point_eval
import pyoracle
import sqd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("key", ec.KEYS)
def test_device_intervals_enclose_point_values(key):
    prog = sqd.Program.from_lemma(key)
    text = prog.text()
    pp = point_eval.Program(text)
    o = pyoracle.Oracle(text)
    (lb, ub), rng = ec.oracle_boxes(o, text)
    with sqd.Core(prog) as core:
        glb, gub = core.initial_box()                 # the oracle's boxes are the device's
        assert np.array_equal(glb.view(np.uint64), lb[:, 0].view(np.uint64))
        evaluations = {prec: core.eval_program(lb, ub, prec) for prec in (64, 32)}
    assert evaluations[64][0].shape[1] == len(pp.bounds) and evaluations[64][1].shape[1] == len(pp.constraints)
    for prec, (b, c) in evaluations.items():            # bit for bit the oracle's evaluation
        ob, oc = o.eval_program(lb, ub, len(pp.bounds), len(pp.constraints), prec)
        assert np.array_equal(b.view(np.uint64), ob.view(np.uint64)), (key, prec)
        assert np.array_equal(c, oc), (key, prec)
    ec.check(key, pp, lb, ub, evaluations, rng)
