"""The offline input tools as command lines (host/tools_main.cpp over
libg2host.so): preprocess and kcore byte-identical to the reference tools'
outputs (tests/golden, written by the tools compiled from kmeans/preprocess.cpp
and kmeans/kcore.cpp); k-means view binning and view packing pinned by
planted communities, determinism and the invariants of the packing."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "g2-aimd_amd")


def tool(name, *args):
    p = subprocess.run([os.path.join(BIN, name)] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


@pytest.mark.parametrize("k", [0, 1, 2])
def test_preprocess_and_kcore_byte_identical(tmp_path, k):
    txt = tmp_path / "g.txt"
    shutil.copy(os.path.join(GOLD, f"toy_{k}.txt"), txt)
    tool("preprocess", "-f", txt)
    assert (tmp_path / "g.bin").read_bytes() == open(os.path.join(GOLD, f"toy_{k}.pre.bin"), "rb").read()
    tool("kcore", "-f", tmp_path / "g.bin", "-nb", 4)
    assert (tmp_path / "g.bin").read_bytes() == open(os.path.join(GOLD, f"toy_{k}.kcore.bin"), "rb").read()
    assert (tmp_path / "g.bin.2.hop.vbmap").read_bytes() == \
        open(os.path.join(GOLD, f"toy_{k}.kcore.vbmap"), "rb").read()


def planted(g2, sizes, p_in, p_out, seed):
    """disjoint dense communities with a few random edges between them"""
    rng = np.random.default_rng(seed)
    pairs, off, label = [], 0, []
    for c, s in enumerate(sizes):
        for i in range(s):
            for j in range(i + 1, s):
                if rng.random() < p_in:
                    pairs.append((off + i, off + j))
        label += [c] * s
        off += s
    for _ in range(int(p_out * off)):
        a, b = rng.integers(0, off, 2)
        if a != b:
            pairs.append((int(a), int(b)))
    return G.from_pairs(g2, pairs, off), np.array(label)


def kmeans(g2, g, k1, k2=1, iters=20, seed=7):
    import ctypes as C
    L = g2.host_lib()
    u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
    L.g2h_kmeans.argtypes = [u64p, C.c_uint64, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, u32p,
                             C.POINTER(C.c_uint32)]
    out = np.zeros(g.nv, np.uint32)
    n = C.c_uint32()
    assert L.g2h_kmeans(g.row_ptrs, g.nv, g.cols, k1, k2, iters, seed, out, C.byref(n)) == 0
    return out, n.value


def view_pack(g2, g, cluster, hop, max_entries, pool=1, by_degree=0):
    import ctypes as C
    L = g2.host_lib()
    u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
    L.g2h_view_pack.argtypes = [u64p, C.c_uint64, u32p, u32p, C.c_uint32, C.c_int, C.c_uint64, C.c_uint32,
                                C.c_int, u32p, C.POINTER(C.c_uint32)]
    bins = np.zeros(g.nv, np.uint32)
    nb = C.c_uint32()
    assert L.g2h_view_pack(g.row_ptrs, g.nv, g.cols, np.ascontiguousarray(cluster, np.uint32), int(cluster.max()) + 1,
                           hop, max_entries, pool, by_degree, bins, C.byref(nb)) == 0
    return bins, nb.value


def same_partition(a, b):
    """equal up to renaming the parts"""
    m = {}
    for x, y in zip(a.tolist(), b.tolist()):
        if m.setdefault(x, y) != y:
            return False
    return len(set(m.values())) == len(m)


def test_kmeans_recovers_planted_communities(g2):
    g, label = planted(g2, [60, 80, 70, 90], 0.5, 0.05, 3)
    cl, n = kmeans(g2, g, 4)
    assert n == 4 and same_partition(cl, label)
    # the same seed gives the same clusters; two levels give k1 * k2 at most
    assert np.array_equal(kmeans(g2, g, 4)[0], cl)
    cl2, n2 = kmeans(g2, g, 2, 2)
    assert 2 <= n2 <= 4 and cl2.max() == n2 - 1


def test_view_pack_invariants(g2, orc):
    g = g2.gen_rmat(11, 16, 0.57, 0.19, 0.19, 0x6732A1D0C0000002).preprocess()
    cl, n = kmeans(g2, g, 4, iters=10)
    # a limit above the whole graph: each cluster packs into one bin, then all merge
    bins, nb = view_pack(g2, g, cl, 2, 10 * g.ne)
    assert nb == 1 and not bins.any()
    # a tight limit: every bin's hop-2 closure stays within it unless one source fills a bin
    limit = g.ne // 6
    bins, nb = view_pack(g2, g, cl, 2, limit, pool=2)
    assert nb > 1 and set(bins.tolist()) == set(range(nb))
    for b in range(nb):
        src = np.nonzero(bins == b)[0].astype(np.uint32)
        view = g2.view_materialize(g, src, 2)
        assert view.ne <= limit or len(src) == 1
    # matches are a sum over bins whatever the packing (oracle, 4-cliques)
    p = orc.OrcPlan.from_buffer_copy(bytes(g2.plan_for_preset(24)))
    whole = orc.gm(g, p)["count"]
    assert sum(orc.gm(g, p, sources=np.nonzero(bins == b)[0].astype(np.uint32))["count"] for b in range(nb)) == whole


def test_kmeans_and_view_pack_cli(g2, tmp_path):
    g, _ = planted(g2, [50, 50, 50, 50], 0.5, 0.05, 4)
    path = tmp_path / "p.bin"
    g.write_bin(str(path))
    tool("kmeans", "-f", path, "-k", "4", "-i", "10")
    part = (tmp_path / "p.bin.kmeans.4").read_text().split()
    assert len(part) == g.nv and set(part) == {"0", "1", "2", "3"}
    tool("view_pack", "-gf", path, "-pf", tmp_path / "p.bin.kmeans.4", "-h", "2", "-m", "1e-5", "-s", "1")
    lines = (tmp_path / "p.bin.2.hop.vbmap").read_text().splitlines()
    nb = int(lines[0])
    assert len(lines) == g.nv + 1 and {int(x) for x in lines[1:]} == set(range(nb))
    nb2, bins = g2.read_vbmap(str(tmp_path / "p.bin.2.hop.vbmap"), g.nv)
    assert nb2 == nb and bins.max() == nb - 1


@pytest.mark.parametrize("tool_name,exts", [("toGAMMA", ("col", "dst", "vlabel")), ("toGSI", ("g",)),
                                            ("toPangolin", ("pangolin.txt",))])
def test_converters_byte_identical(tmp_path, tool_name, exts):
    """toGAMMA / toGSI / toPangolin outputs equal the reference converters'
    (tests/golden/make_converter_fixtures.py)"""
    g = tmp_path / "toy_0.bin"
    shutil.copy(os.path.join(GOLD, "toy_0.pre.bin"), g)
    tool(tool_name, g)
    for ext in exts:
        assert (tmp_path / f"toy_0.{ext}").read_bytes() == open(os.path.join(GOLD, f"toy_0.{ext}"), "rb").read(), ext
