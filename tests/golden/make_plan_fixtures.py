"""Write tests/golden/plans.json: the reference's own plan tables.

Runs oracle/_ref/plan_dump -- the reference's Graph (presets, ReMapVertexId,
bliss conditions) and Plan code (system/plan.h), compiled from the sources in
/root/reference by `make -C oracle ref` -- on every preset pattern and on
seeded random connected patterns, and stores the printed tables.  Needs the
reference checkout, so it runs in the build container only; the JSON it
writes is the fixture the tests read.
"""
import importlib
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
g2 = importlib.import_module("g2-aimd_amd")

PRESETS = list(range(0, 19)) + [23, 24, 25, 26, 27]


def edges_of(qadj):
    n = len(qadj)
    return [(a, b) for a in range(n) for b in range(a + 1, n) if qadj[a] >> b & 1]


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "plan_dump")
    out = {"presets": {}, "random": []}
    for q in PRESETS:
        r = subprocess.run([exe, "-q", str(q)], check=True, capture_output=True, text=True)
        out["presets"][str(q)] = json.loads(r.stdout)
    # random connected patterns of 3..7 vertices (same generator as the
    # batched-matching workload: density uniform in [0.4, 1.0])
    for j in range(48):
        n = 3 + j % 5
        dens = 0.4 + 0.6 * ((j * 7919) % 97) / 96.0
        qadj = g2.query_random(n, dens, 40000 + j)
        e = edges_of(qadj)
        args = [exe, "-e", str(n)] + [str(x) for ab in e for x in ab]
        r = subprocess.run(args, check=True, capture_output=True, text=True)
        out["random"].append({"seed": 40000 + j, "n": n, "density": dens, "qadj": qadj,
                              "plan": json.loads(r.stdout)})
    with open(os.path.join(HERE, "plans.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", len(out["presets"]), "presets and", len(out["random"]), "random plans")


if __name__ == "__main__":
    main()
