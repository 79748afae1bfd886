"""Write tests/golden/headline_counts.json: oracle totals of the full-size
workloads, so GPU tests and bench.py can check their counts without running
the CPU restatement for an hour.

Each entry names a seeded graph from the project's generators (the same call
bench.py and the tests make), a fingerprint of the generated CSR (V, E and a
position-weighted checksum of cols, so a changed generator is caught), and the
oracle's results on it:
  rmat-<s>        triangles (q=0) and 4-cliques (q=24) from oracle.gm (the
                  restatement of app_gmatch/gmatch.h), with orc_kclique_count
                  (an independent ordered-merge k-clique counter) beside them
  social-<n>      maximal-clique total and size histogram from oracle.mce (the
                  restatement of app_BK/BK.h), roots in degeneracy order

Usage:  python tests/golden/make_headline_fixtures.py rmat:22 social:2000000 ...
Entries already in the file are kept unless named again.  The R-MAT 22 entry
takes about 75 minutes on 8 cores.
"""
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, "headline_counts.json")

RMAT_SEED = 0x6732A1D0C0000002
SOCIAL_SEED = 0x6732A1D0C0000003


def fingerprint(g, chunk=1 << 26):
    """(V, E, sum over i of cols[i] * (i + 1) mod 2^64) of a CSR graph, in
    chunks (a 65M-vertex graph has 3.6e9 entries)."""
    chk = 0
    for a in range(0, g.ne, chunk):
        b = min(a + chunk, g.ne)
        w = np.arange(a + 1, b + 1, dtype=np.uint64)
        chk = (chk + int(np.sum(g.cols[a:b].astype(np.uint64) * w, dtype=np.uint64))) % (1 << 64)
    return {"nv": int(g.nv), "ne": int(g.ne), "cols_checksum": chk}


def rmat_graph(g2, scale):
    return g2.gen_rmat(scale, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess()


def social_graph(g2, n):
    """Same parameters as bench.py's social workload at n vertices."""
    f = n / 65e6
    g = g2.gen_social(n, 2.1, 2.0, 5000.0, int(1.54e9 * f), int(2e6 * f), 10, 200, 3.0, 0.3, 0.9, 1e4,
                      SOCIAL_SEED)
    g = g.preprocess()
    g, _, _, _ = g.kcore_recode()
    return g


def rmat_entry(g2, orc, scale, threads):
    g = rmat_graph(g2, scale)
    e = {"graph": f"R-MAT scale {scale}, edge factor 16, (.57,.19,.19,.05), seed {RMAT_SEED:#x}, "
                  "preprocessed", **fingerprint(g)}
    for q in (0, 24):
        t = time.time()
        p = orc.OrcPlan.from_buffer_copy(bytes(g2.plan_for_preset(q)))
        e[f"q{q}"] = orc.gm(g, p, threads=threads)["count"]
        e[f"q{q}_seconds"] = round(time.time() - t, 1)
        print(f"rmat-{scale} q={q}: {e[f'q{q}']} in {e[f'q{q}_seconds']}s", flush=True)
    t = time.time()
    e["kclique3"] = orc.kclique_count(g, 3, threads=threads)
    print(f"rmat-{scale} k=3 (independent): {e['kclique3']} in {time.time() - t:.1f}s", flush=True)
    assert e["kclique3"] == e["q0"], "the two oracles disagree on triangles"
    if scale <= 20:
        e["kclique4"] = orc.kclique_count(g, 4, threads=threads)
        assert e["kclique4"] == e["q24"], "the two oracles disagree on 4-cliques"
    return e


def social_entry(g2, orc, n, threads):
    g = social_graph(g2, n)
    fp = fingerprint(g)
    print(f"social-{n}: {fp}", flush=True)
    t = time.time()
    r = orc.mce(g, threads=threads, hist_len=256)
    hist = {int(k): int(v) for k, v in enumerate(r["hist"]) if v}
    e = {"graph": f"social generator n={n}, seed {SOCIAL_SEED:#x}, preprocessed, degeneracy order",
         **fp, "max_cliques": r["total"], "hist": hist, "nodes": r["nodes"],
         "seconds": round(time.time() - t, 1)}
    print(f"social-{n}: {r['total']} maximal cliques in {e['seconds']}s", flush=True)
    return e


def main(argv):
    g2 = importlib.import_module("g2-aimd_amd")
    import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for spec in argv:
        kind, size = spec.split(":")
        if kind == "rmat":
            data[f"rmat-{size}"] = rmat_entry(g2, orc, int(size), threads)
        elif kind == "social":
            data[f"social-{size}"] = social_entry(g2, orc, int(size), threads)
        else:
            raise SystemExit(f"unknown entry {spec}")
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main(sys.argv[1:] or ["rmat:17", "rmat:22"])
