"""oracle -- ctypes binding of the CPU restatement (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, to check results.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
Q = 8


class OrcPlan(C.Structure):
    """Same layout as g2_plan (include/g2plan.h)."""

    _fields_ = [
        ("n", C.c_uint32), ("root_degree", C.c_uint32), ("hop", C.c_int32),
        ("seq", C.c_uint32 * Q), ("rev", C.c_uint32 * Q), ("bn_count", C.c_uint32 * Q),
        ("bn", (C.c_uint32 * Q) * Q), ("cond_count", C.c_uint32 * Q),
        ("cond", ((C.c_uint32 * 2) * Q) * Q), ("shared", C.c_uint32 * Q),
        ("pre_bn_count", C.c_uint32 * Q), ("pre_bn", (C.c_uint32 * Q) * Q),
        ("after_bn_count", C.c_uint32 * Q), ("after_bn", (C.c_uint32 * Q) * Q),
        ("pre_cond_count", C.c_uint32 * Q), ("pre_cond", ((C.c_uint32 * 2) * Q) * Q),
        ("after_cond_count", C.c_uint32 * Q), ("after_cond", ((C.c_uint32 * 2) * Q) * Q),
        ("strategy", C.c_uint32 * (Q + 1)), ("moving_lvl", C.c_uint32 * Q),
    ]

    @staticmethod
    def from_dict(d: dict) -> "OrcPlan":
        """Build from plan tables in the reference's shapes (plan_dump JSON)."""
        p = OrcPlan()
        n = d["n"]
        p.n, p.root_degree, p.hop = n, d["root_degree"], d["hop"]
        for i in range(n):
            p.seq[i] = d["seq"][i]
            p.rev[i] = d["rev"][i]
            p.shared[i] = d["shared"][i]
            p.moving_lvl[i] = d["moving_lvl"][i]
            for key in ("bn", "pre_bn", "after_bn"):
                getattr(p, key + "_count")[i] = len(d[key][i])
                for j, x in enumerate(d[key][i]):
                    getattr(p, key)[i][j] = x
            for key in ("cond", "pre_cond", "after_cond"):
                getattr(p, key + "_count")[i] = len(d[key][i])
                for k, (op, x) in enumerate(d[key][i]):
                    getattr(p, key)[i][k][0] = op
                    getattr(p, key)[i][k][1] = x
        for i in range(n + 1):
            p.strategy[i] = d["strategy"][i]
        return p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    _lib = bind(path)
    return _lib


def bind(path: str) -> C.CDLL:
    """Load an oracle library built from oracle.cpp and declare its signatures.
    The tests also use it for copies of the oracle with a planted mistake, to
    show that their checks catch it."""
    L = C.CDLL(path)
    P64 = C.POINTER(C.c_uint64)
    L.orc_mce.argtypes = [u64p, C.c_uint64, u32p, C.c_uint64, u32p, C.c_uint64, C.c_int,
                          P64, u64p, C.c_uint32, C.c_void_p, C.c_void_p, P64, P64]
    L.orc_gm.argtypes = [u64p, C.c_uint64, u32p, C.c_uint64, u32p, C.c_uint64, C.POINTER(OrcPlan),
                         C.c_int, P64, C.c_void_p, P64]
    L.orc_bruteforce_mce.argtypes = [u64p, C.c_uint64, u32p, P64, u64p, C.c_uint32]
    L.orc_bruteforce_pattern.argtypes = [u64p, C.c_uint64, u32p, C.c_uint32, u32p, C.c_int, P64]
    L.orc_automorphism_count.argtypes = [C.c_uint32, u32p]
    L.orc_automorphism_count.restype = C.c_uint64
    L.orc_kclique_count.argtypes = [u64p, C.c_uint64, u32p, C.c_uint32, C.c_int, P64]
    L.orc_kclique_count_ranked.argtypes = [u64p, C.c_uint64, u32p, C.c_uint32, u32p, C.c_uint64,
                                           C.c_int, P64]
    L.orc_local_bytes.argtypes = [u64p, C.c_uint64, u32p, u32p, C.c_uint64, C.c_uint32, u64p]
    i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    for f in (L.orc_two_hop, L.orc_two_hop_exact):
        f.argtypes = [u64p, C.c_uint64, u32p, u32p, C.c_uint64, i64p]
    return L


def _cols(g):
    return g.cols if len(g.cols) else np.zeros(1, np.uint32)


def _src(sources, nv):
    if sources is None:
        return np.arange(nv, dtype=np.uint32)
    return np.ascontiguousarray(sources, dtype=np.uint32)


def mce(g, sources=None, hist_len: int = 256, threads: int = 0, per_root: bool = False):
    """Oracle MCE: dict(total, hist, nodes, bytes[, per_root_nodes, per_root_bytes])."""
    src = _src(sources, g.nv)
    hist = np.zeros(hist_len, np.uint64)
    total, nodes, nbytes = C.c_uint64(), C.c_uint64(), C.c_uint64()
    prn = np.zeros(max(len(src), 1), np.uint64) if per_root else None
    prb = np.zeros(max(len(src), 1), np.uint64) if per_root else None
    s = src if len(src) else np.zeros(1, np.uint32)
    rc = lib().orc_mce(g.row_ptrs, g.nv, _cols(g), g.ne, s, len(src), threads, C.byref(total),
                       hist, hist_len, prn.ctypes.data if per_root else None,
                       prb.ctypes.data if per_root else None, C.byref(nodes), C.byref(nbytes))
    assert rc == 0
    out = {"total": total.value, "hist": hist, "nodes": nodes.value, "bytes": nbytes.value}
    if per_root:
        out["per_root_nodes"] = prn[: len(src)]
        out["per_root_bytes"] = prb[: len(src)]
    return out


def gm(g, plan: OrcPlan, sources=None, threads: int = 0, per_root: bool = False):
    src = _src(sources, g.nv)
    cnt, nbytes = C.c_uint64(), C.c_uint64()
    prc = np.zeros(max(len(src), 1), np.uint64) if per_root else None
    s = src if len(src) else np.zeros(1, np.uint32)
    rc = lib().orc_gm(g.row_ptrs, g.nv, _cols(g), g.ne, s, len(src), C.byref(plan), threads,
                      C.byref(cnt), prc.ctypes.data if per_root else None, C.byref(nbytes))
    assert rc == 0, rc
    out = {"count": cnt.value, "bytes": nbytes.value}
    if per_root:
        out["per_root_count"] = prc[: len(src)]
    return out


def bruteforce_mce(g, hist_len: int = 64):
    hist = np.zeros(hist_len, np.uint64)
    total = C.c_uint64()
    rc = lib().orc_bruteforce_mce(g.row_ptrs, g.nv, _cols(g), C.byref(total), hist, hist_len)
    assert rc == 0
    return total.value, hist


def bruteforce_pattern(g, qadj, threads: int = 0) -> int:
    """Injective homomorphisms of a pattern whose vertex order is BFS-connected."""
    q = np.array(list(qadj) + [0] * (Q - len(qadj)), np.uint32)
    out = C.c_uint64()
    rc = lib().orc_bruteforce_pattern(g.row_ptrs, g.nv, _cols(g), len(qadj), q, threads,
                                      C.byref(out))
    assert rc == 0
    return out.value


def automorphism_count(qadj) -> int:
    q = np.array(list(qadj) + [0] * (Q - len(qadj)), np.uint32)
    return int(lib().orc_automorphism_count(len(qadj), q))


def kclique_count(g, k: int, threads: int = 0) -> int:
    out = C.c_uint64()
    rc = lib().orc_kclique_count(g.row_ptrs, g.nv, _cols(g), k, threads, C.byref(out))
    assert rc == 0
    return out.value


def kclique_count_ranked(g, k: int, sources=None, threads: int = 0) -> int:
    """k-cliques whose lowest vertex under (degree, id) is a source."""
    src = _src(sources, g.nv)
    s = src if len(src) else np.zeros(1, np.uint32)
    out = C.c_uint64()
    rc = lib().orc_kclique_count_ranked(g.row_ptrs, g.nv, _cols(g), k, s, len(src), threads, C.byref(out))
    assert rc == 0, rc
    return out.value


def local_bytes(g, n: int, sources=None) -> np.ndarray:
    """per-pass algorithmic bytes of the local-bitmap matching kernels (u64[5])"""
    src = _src(sources, g.nv)
    s = src if len(src) else np.zeros(1, np.uint32)
    out = np.zeros(5, np.uint64)
    rc = lib().orc_local_bytes(g.row_ptrs, g.nv, _cols(g), s, len(src), n, out)
    assert rc == 0, rc
    return out


def two_hop(g, sources=None) -> np.ndarray:
    """HyperLogLog estimate (64 registers) of the 2-hop ball size |{v} u N(v) u N(N(v))|
    per source, the integers the device feature kernel must reproduce."""
    src = _src(sources, g.nv)
    out = np.zeros(max(len(src), 1), np.int64)
    s = src if len(src) else np.zeros(1, np.uint32)
    assert lib().orc_two_hop(g.row_ptrs, g.nv, _cols(g), s, len(src), out) == 0
    return out[: len(src)]


def two_hop_exact(g, sources=None) -> np.ndarray:
    """exact 2-hop ball size per source (brute force)"""
    src = _src(sources, g.nv)
    out = np.zeros(max(len(src), 1), np.int64)
    s = src if len(src) else np.zeros(1, np.uint32)
    assert lib().orc_two_hop_exact(g.row_ptrs, g.nv, _cols(g), s, len(src), out) == 0
    return out[: len(src)]
