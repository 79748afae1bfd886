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
