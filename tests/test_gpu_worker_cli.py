"""The `bk` and `run` command lines (host/worker.cpp: the reference's
system/worker.h:28-280 producer/consumer pipeline over view bins, with its
flags and output lines) against the oracle's whole-graph results.

The graph is cut into 4 kcore bins, materialised as hop-2 views by the
producers and counted by 2 consumers sharing the GPU; the printed and JSON
totals must equal the oracle's, since every count is a sum over roots."""
import json
import os
import subprocess
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "g2-aimd_amd")
SEED = 0x6732A1D0C0000011


def run_cli(name, args, out_json, timeout=90):
    """Run a command line, echoing its output as it comes (long runs keep
    printing), killed after `timeout` seconds."""
    cmd = [os.path.join(BIN, name)] + args + ["-json", str(out_json)]
    print("case:", " ".join(cmd), flush=True)
    lines = []
    with subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          errors="replace") as p:
        timer = threading.Timer(timeout, p.kill)
        timer.start()
        try:
            for line in p.stdout:
                lines.append(line)
                print("  |", line, end="", flush=True)
            rc = p.wait()
        finally:
            timer.cancel()
    log = "".join(lines)
    assert rc == 0, f"rc={rc}\n" + log[-3000:]
    with open(out_json) as f:
        return json.load(f), log


@pytest.fixture(scope="module")
def binned_graph(g2, tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    g = g2.gen_chung_lu(20000, 2.3, 2.0, 500.0, 150000, SEED).preprocess()
    g, _, shells, _ = g.kcore_recode()
    path = str(d / "cl.bin")
    g.write_bin(path)
    g2.write_vbmap(path + ".2.hop.vbmap", 4, g2.kcore_bins(shells, 4, g.nv))
    return g, path, d


def test_bk_matches_oracle(orc, binned_graph):
    g, path, d = binned_graph
    r, log = run_cli("bk", ["-dg", path, "-h", "2", "-pn", "2", "-cn", "2", "-t", "2", "-qs", "2",
                            "-ss", "1"], d / "bk.json")
    ref = orc.mce(g, hist_len=128)
    assert r["app"] == "bk" and r["total"] == ref["total"] > 0
    assert {int(k): v for k, v in r["hist"].items()} == \
        {int(k): int(ref["hist"][k]) for k in np.nonzero(ref["hist"])[0]}
    assert f"Total counts: {ref['total']}" in log


def test_bk_reordered_and_split_bins(orc, binned_graph):
    """-dr 1 orders the bins by their load factor and -ds 1 splits the bins
    furthest above the mean until there are (4/2 + 1) x 2 = 6
    (view/view_bin_manager.h's Reorder and Split): the sources are still
    partitioned, so the total and histogram stay the oracle's, over 6 views."""
    g, path, d = binned_graph
    r, log = run_cli("bk", ["-dg", path, "-h", "2", "-pn", "2", "-cn", "2", "-t", "2", "-qs", "2",
                            "-ss", "1", "-dr", "1", "-ds", "1"], d / "bk_split.json")
    ref = orc.mce(g, hist_len=128)
    assert r["total"] == ref["total"] > 0
    assert {int(k): v for k, v in r["hist"].items()} == \
        {int(k): int(ref["hist"][k]) for k in np.nonzero(ref["hist"])[0]}
    assert r["plain"]["launches"] == 6


def test_bk_reverse_schedule(orc, binned_graph):
    """-sched reverse takes each view's sources backwards (the roots with the
    largest neighbourhoods first) through g2_sched.order, with no cost model:
    the order of the roots changes, the total and histogram do not."""
    g, path, d = binned_graph
    r, log = run_cli("bk", ["-dg", path, "-h", "2", "-pn", "2", "-cn", "2", "-t", "2", "-qs", "2",
                            "-ss", "1", "-sched", "reverse"], d / "bk_reverse.json")
    ref = orc.mce(g, hist_len=128)
    assert r["total"] == ref["total"] > 0
    assert {int(k): v for k, v in r["hist"].items()} == \
        {int(k): int(ref["hist"][k]) for k in np.nonzero(ref["hist"])[0]}
    assert "reverse" in r and "plain" not in r
    assert r["reverse"]["launches"] == 4


@pytest.mark.parametrize("q", [0, 24])
def test_run_matches_oracle(g2, orc, binned_graph, q):
    g, path, d = binned_graph
    p = g2.plan_for_preset(q)
    assert p.hop == 2
    r, log = run_cli("run", ["-dg", path, "-q", str(q), "-pn", "2", "-cn", "2", "-qs", "2"],
                     d / f"run{q}.json")
    ref = orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(p)))["count"]
    assert r["app"] == "run" and r["query"] == q and r["total"] == ref > 0
    assert f"Total counts: {ref}" in log


def fixture_entry(name):
    with open(os.path.join(ROOT, "tests", "golden", "headline_counts.json")) as f:
        return json.load(f)[name]


def test_bk_learned_social_pipeline(g2, tmp_path):
    """The social MCE workload through the drop-in pipeline: 4 kcore bins,
    4 producers, 4 consumers, queue 4, sorted sources; the plain schedule and
    the learned + batched one (cost model trained by bk itself on two other
    seeds) both give the oracle's total and histogram (fixture)."""
    fx = fixture_entry("social-1000000")
    model = tmp_path / "social.model"
    r, log = run_cli("bk", ["-gen", "social:1000000", "-h", "2", "-pn", "4", "-cn", "4", "-qs", "4", "-ss", "1",
                            "-sched", "both", "-train", str(model),
                            "-train-gen", "social:500000:0x6732A1D0C0000021,social:500000:0x6732A1D0C0000022",
                            "-epochs", "10", "-progress", "1"], tmp_path / "social.json", timeout=400)
    assert r["vertices"] == fx["nv"] and 2 * r["edges"] == fx["ne"]
    assert r["total"] == fx["max_cliques"]
    assert {int(k): v for k, v in r["hist"].items()} == {int(k): v for k, v in fx["hist"].items()}
    for form in ("plain", "learned"):
        assert r[form]["total"] == fx["max_cliques"]
        assert r[form]["hist"] == r["hist"]
    assert r["learned"]["split_roots"] >= 0 and r["plain"]["donations"] == 0
    assert model.exists() and "learned vs plain" in log
    print(f"plain {r['plain']['seconds']:.3f}s kernel {r['plain']['kernel_ms']:.1f} ms idle "
          f"{r['plain']['wave_idle_frac']:.3f}; learned {r['learned']['seconds']:.3f}s kernel "
          f"{r['learned']['kernel_ms']:.1f} ms idle {r['learned']['wave_idle_frac']:.3f}", flush=True)
    # the model file is reusable (on a training graph: other small seeds can
    # hold a dense community with 1e9+ maximal cliques)
    r2, _ = run_cli("bk", ["-gen", "social:500000:0x6732A1D0C0000021", "-h", "2", "-pn", "2", "-cn", "2",
                           "-qs", "2", "-sched", "both", "-model", str(model)], tmp_path / "again.json", timeout=200)
    assert r2["plain"]["total"] == r2["learned"]["total"] > 0
    assert r2["plain"]["hist"] == r2["learned"]["hist"]


def test_bk_aimd_retry(orc, g2, tmp_path):
    """A budget that holds everything but the whole view's deferred-root list:
    the first launch overflows, the AIMD chunk loop (system/subgraph_container.h:
    178-211) retries smaller chunks, and the total stays the oracle's."""
    args = ["-gen", "chunglu:100000", "-nb", "1", "-h", "2", "-pn", "1", "-cn", "1", "-qs", "1"]
    r, _ = run_cli("bk", args, tmp_path / "free.json")
    g = g2.gen_chung_lu(100000, 2.3, 2.0, 2000.0, 1000000, 0x6732A1D0C0000001).preprocess()
    g, _, _, _ = g.kcore_recode()
    ref = orc.mce(g, hist_len=128)
    assert r["total"] == ref["total"] and r["plain"]["retries"] == 0
    # the deferred-root list (8 B per root, 256-B granules) is the last buffer a run takes
    budget = r["plain"]["device_bytes"] - (8 * g.nv + 255) // 256 * 256 + 8 * 20000
    r2, log = run_cli("bk", args + ["-budget", repr(budget / 1e9)], tmp_path / "tight.json")
    assert r2["plain"]["retries"] >= 1 and "Iteration failed!!" in log and "--- chunk: " in log
    assert r2["total"] == ref["total"]
    assert {int(k): v for k, v in r2["hist"].items()} == \
        {int(k): int(ref["hist"][k]) for k in np.nonzero(ref["hist"])[0]}


def test_run_pattern_file(g2, orc, binned_graph):
    """run -qf: a pattern given as an edge list instead of a preset; the total
    equals the oracle's count of the same plan (and the brute-force count of
    injective homomorphisms over |Aut| on the whole graph)."""
    g, path, d = binned_graph
    qadj = g2.query_random(5, 0.7, 40021)
    pf = d / "pattern.txt"
    with open(pf, "w") as f:
        f.write("# a random connected 5-vertex pattern\n")
        for a in range(5):
            for b in range(a + 1, 5):
                if qadj[a] >> b & 1:
                    f.write(f"{a} {b}\n")
    r, log = run_cli("run", ["-dg", path, "-qf", str(pf), "-pn", "2", "-cn", "2", "-qs", "2"], d / "runqf.json")
    p = g2.plan_build(g2.query_remap(qadj))
    ref = orc.gm(g, orc.OrcPlan.from_buffer_copy(bytes(p)))["count"]
    assert r["total"] == ref
    assert f"Total counts: {ref}" in log
