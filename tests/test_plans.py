"""The product's plan builder (libg2host.so, restating system/plan.h and the
query side of common/graph.h) against the reference's own tables, printed by
the reference's code compiled from its sources (tests/golden/plans.json)."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "plans.json")) as f:
    FIX = json.load(f)


def rows(qadj):
    n = len(qadj)
    return [[v for v in range(n) if qadj[u] >> v & 1] for u in range(n)]


@pytest.mark.parametrize("q", sorted(FIX["presets"], key=int))
def test_preset_plans(g2, q):
    ref = FIX["presets"][q]
    n, qadj = g2.query_preset(int(q))
    assert rows(qadj) == ref["rows"]
    assert [list(c) for c in g2.query_conditions(qadj)] == ref["conditions"]
    d = g2.plan_build(qadj).to_dict()
    for k, v in d.items():
        assert v == ref[k], k


def test_random_plans(g2):
    for item in FIX["random"]:
        ref = item["plan"]
        qadj = g2.query_remap(item["qadj"])
        assert rows(qadj) == ref["rows"], item["seed"]
        assert [list(c) for c in g2.query_conditions(qadj)] == ref["conditions"], item["seed"]
        d = g2.plan_build(qadj).to_dict()
        for k, v in d.items():
            assert v == ref[k], (item["seed"], k)


def test_random_pattern_generator(g2):
    for j, item in enumerate(FIX["random"]):
        assert g2.query_random(item["n"], item["density"], item["seed"]) == item["qadj"]


def test_disconnected_pattern_rejected(g2):
    with pytest.raises(RuntimeError):
        g2.plan_build([0b10, 0b01, 0b1000, 0b0100])


def test_malformed_patterns_rejected(g2):
    """Adjacency masks must be symmetric, without self loops or bits at or
    above n; the plan and condition builders refuse anything else, and the
    automorphism count and canonical code are 0 for it."""
    for bad in ([0b10, 0b00, 0b00],          # 0-1 listed on one side only
                [0b11, 0b01],                # self loop at 0
                [0b1010, 0b0001, 0b0000]):   # bit 3 with n = 3
        with pytest.raises(RuntimeError, match="symmetric"):
            g2.plan_build(bad)
        with pytest.raises(RuntimeError, match="symmetric"):
            g2.query_conditions(bad)
        with pytest.raises(RuntimeError, match="symmetric"):
            g2.query_remap(bad)
        assert g2.query_automorphisms(bad) == 0 and g2.query_canonical(bad) == 0
    assert g2.query_automorphisms([0b110, 0b101, 0b011]) == 6
