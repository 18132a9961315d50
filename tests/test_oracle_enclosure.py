"""Lemma-level enclosure of the oracle's interval evaluation (the checker the
device is compared with bit for bit): every bound expression and constraint of
each statement on boxes the proofs visit, in binary64 and binary32, against the
real values at sampled points (checks in tests/enclosure_check.py)."""
import pytest

import enclosure_check as ec
import point_eval
import pyoracle
import sqd


@pytest.mark.parametrize("key", ec.KEYS)
def test_oracle_intervals_enclose_point_values(key):
    text = sqd.Program.from_lemma(key).text()
    pp = point_eval.Program(text)
    o = pyoracle.Oracle(text)
    (lb, ub), rng = ec.oracle_boxes(o, text)
    evaluations = {prec: o.eval_program(lb, ub, len(pp.bounds), len(pp.constraints), prec) for prec in (64, 32)}
    ec.check(key, pp, lb, ub, evaluations, rng)


def test_check_catches_an_interval_that_misses_points():
    # a planted mistake: one bound's intervals shrunk to their upper end and one
    # constraint claimed definitely true must fail the check
    text = sqd.Program.from_lemma("one").text()
    pp = point_eval.Program(text)
    o = pyoracle.Oracle(text)
    (lb, ub), rng = ec.oracle_boxes(o, text)
    b, c = o.eval_program(lb, ub, len(pp.bounds), len(pp.constraints), 64)
    shrunk = b.copy()
    shrunk[:, 3, 0] = shrunk[:, 3, 1]
    with pytest.raises(AssertionError):
        ec.check("one", pp, lb, ub, {64: (shrunk, c)}, rng)
    decided = c.copy()
    decided[:, :, 0] = 1
    with pytest.raises(AssertionError):
        ec.check("one", pp, lb, ub, {64: (b, decided)}, ec.oracle_boxes(o, text)[1])
