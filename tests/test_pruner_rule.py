"""The pruner's fallback rule: per-QP partition-head thresholds, parsed from
the command-line forms bench.py and pcme.eval_pruner accept (host logic)."""
import pytest

from pcme import pruner as P


def test_default_rule_is_per_qp():
    assert {q: P.thr_part_for(q) for q in (22, 27, 32, 37)} == {22: 0.7, 27: 0.7, 32: 0.8, 37: 0.9}
    # a QP without an entry of its own keeps the base threshold
    assert P.thr_part_for(30) == 0.7


@pytest.mark.parametrize("text,want", [
    (None, {22: 0.7, 27: 0.7, 32: 0.8, 37: 0.9}),
    ("", {22: 0.7, 27: 0.7, 32: 0.8, 37: 0.9}),
    ("0.6", {22: 0.6, 27: 0.6, 32: 0.6, 37: 0.6}),
    ("22:0.75,27:0.7,32:0.85,37:0.95", {22: 0.75, 27: 0.7, 32: 0.85, 37: 0.95}),
    ("27:0.5", {22: 0.7, 27: 0.5, 32: 0.7, 37: 0.7}),
])
def test_parse_and_lookup(text, want):
    rule = P.parse_thr(text)
    assert {q: P.thr_part_for(q, rule) for q in want} == want
