"""g2-aimd_amd -- MI355X-native subgraph-search hot path of G2-AIMD.

Python mirror of the reference's app interface over the two C ABIs:

* ``libg2host.so`` (include/g2host.h): CSR graphs, generators, kcore recode,
  view bins, query plans (common/graph.h, system/plan.h, view/*.h,
  kmeans/kcore.cpp of the reference);
* ``libg2hip.so`` (include/g2hip.h): the gfx950 kernels behind the
  ``PipelineExecutor`` boundary (system/pipeline_executor.h:44-207):
  ``Executor.transfer`` ~ ``Transfer``, ``Executor.run_mce`` /
  ``Executor.run_gm`` ~ ``Run`` with the BK / GMatch apps.

The package is imported by directory name (it contains a hyphen)::

    g2 = importlib.import_module("g2-aimd_amd")

There is no CPU fallback: every compute call goes through libg2hip.so and
raises if the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
G2_MAXQ = 8

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class G2Plan(C.Structure):
    """ctypes mirror of ``g2_plan`` (include/g2plan.h)."""

    Q = G2_MAXQ
    _fields_ = [
        ("n", C.c_uint32),
        ("root_degree", C.c_uint32),
        ("hop", C.c_int32),
        ("seq", C.c_uint32 * Q),
        ("rev", C.c_uint32 * Q),
        ("bn_count", C.c_uint32 * Q),
        ("bn", (C.c_uint32 * Q) * Q),
        ("cond_count", C.c_uint32 * Q),
        ("cond", ((C.c_uint32 * 2) * Q) * Q),
        ("shared", C.c_uint32 * Q),
        ("pre_bn_count", C.c_uint32 * Q),
        ("pre_bn", (C.c_uint32 * Q) * Q),
        ("after_bn_count", C.c_uint32 * Q),
        ("after_bn", (C.c_uint32 * Q) * Q),
        ("pre_cond_count", C.c_uint32 * Q),
        ("pre_cond", ((C.c_uint32 * 2) * Q) * Q),
        ("after_cond_count", C.c_uint32 * Q),
        ("after_cond", ((C.c_uint32 * 2) * Q) * Q),
        ("strategy", C.c_uint32 * (Q + 1)),
        ("moving_lvl", C.c_uint32 * Q),
    ]

    def to_dict(self) -> dict:
        """Plan tables in the reference's shapes (as oracle/_ref/plan_dump prints)."""
        n = self.n

        def lists(cnt, tab):
            return [[tab[i][j] for j in range(cnt[i])] for i in range(n)]

        def conds(cnt, tab):
            return [[[tab[i][k][0], tab[i][k][1]] for k in range(cnt[i])] for i in range(n)]

        return {
            "n": n,
            "hop": self.hop,
            "root_degree": self.root_degree,
            "seq": list(self.seq[:n]),
            "rev": list(self.rev[:n]),
            "bn_count": list(self.bn_count[:n]),
            "bn": lists(self.bn_count, self.bn),
            "cond_count": list(self.cond_count[:n]),
            "cond": conds(self.cond_count, self.cond),
            "shared": list(self.shared[:n]),
            "pre_bn_count": list(self.pre_bn_count[:n]),
            "pre_bn": lists(self.pre_bn_count, self.pre_bn),
            "after_bn_count": list(self.after_bn_count[:n]),
            "after_bn": lists(self.after_bn_count, self.after_bn),
            "pre_cond_count": list(self.pre_cond_count[:n]),
            "pre_cond": conds(self.pre_cond_count, self.pre_cond),
            "after_cond_count": list(self.after_cond_count[:n]),
            "after_cond": conds(self.after_cond_count, self.after_cond),
            "strategy": list(self.strategy[: n + 1]),
            "moving_lvl": list(self.moving_lvl[:n]),
        }


# --------------------------------------------------------------------------
# libg2host.so
# --------------------------------------------------------------------------
_host = None


def host_lib() -> C.CDLL:
    global _host
    if _host is not None:
        return _host
    path = os.path.join(_HERE, "libg2host.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    vp = C.c_void_p
    pp = C.POINTER(C.c_void_p)
    L.g2h_last_error.restype = C.c_char_p
    L.g2h_set_progress.argtypes = [C.c_int]
    L.g2h_set_progress.restype = None
    L.g2h_graph_read_bin.argtypes = [C.c_char_p, pp]
    L.g2h_graph_write_bin.argtypes = [vp, C.c_char_p]
    L.g2h_graph_from_csr.argtypes = [u64p, C.c_uint64, u32p, C.c_uint64, pp]
    L.g2h_graph_from_edges.argtypes = [u32p, u32p, C.c_uint64, C.c_uint64, pp]
    for f in ("g2h_graph_nv", "g2h_graph_ne", "g2h_graph_max_degree"):
        getattr(L, f).argtypes = [vp]
        getattr(L, f).restype = C.c_uint64
    L.g2h_graph_row_ptrs.argtypes = [vp]
    L.g2h_graph_row_ptrs.restype = C.POINTER(C.c_uint64)
    L.g2h_graph_cols.argtypes = [vp]
    L.g2h_graph_cols.restype = C.POINTER(C.c_uint32)
    L.g2h_graph_free.argtypes = [vp]
    L.g2h_graph_free.restype = None
    L.g2h_graph_preprocess.argtypes = [vp]
    L.g2h_graph_kcore_recode.argtypes = [vp, vp, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
    L.g2h_kcore_bins.argtypes = [u64p, C.c_uint64, C.c_uint32, u32p, C.c_uint64]
    L.g2h_write_vbmap.argtypes = [C.c_char_p, C.c_uint32, u32p, C.c_uint64]
    L.g2h_read_vbmap.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), u32p, C.c_uint64]
    L.g2h_gen_chung_lu.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                   C.c_uint64, pp]
    L.g2h_gen_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                               C.c_uint64, pp]
    L.g2h_gen_social.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                 C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.c_uint64, pp]
    qa = C.c_uint32 * G2_MAXQ
    L.g2h_query_preset.argtypes = [C.c_int, C.POINTER(C.c_uint32), qa]
    L.g2h_query_remap.argtypes = [C.c_uint32, qa, qa]
    L.g2h_query_conditions.argtypes = [C.c_uint32, qa, C.POINTER(C.c_uint32),
                                       C.c_uint32 * (2 * G2_MAXQ * G2_MAXQ)]
    L.g2h_query_automorphisms.argtypes = [C.c_uint32, qa]
    L.g2h_query_automorphisms.restype = C.c_uint64
    L.g2h_query_canonical.argtypes = [C.c_uint32, qa]
    L.g2h_query_canonical.restype = C.c_uint64
    L.g2h_plan_build.argtypes = [C.c_uint32, qa, C.POINTER(G2Plan)]
    L.g2h_query_random.argtypes = [C.c_uint32, C.c_double, C.c_uint64, qa]
    L.g2h_view_materialize.argtypes = [vp, u32p, C.c_uint64, C.c_int, C.c_uint64, u64p, u32p,
                                       C.c_uint64, C.POINTER(C.c_uint64)]
    _host = L
    return L


def _hcheck(rc: int, what: str) -> None:
    if rc != 0:
        msg = host_lib().g2h_last_error().decode()
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


class Graph:
    """CSR data graph (common/graph.h:57-107 data side): ``row_ptrs`` u64[V+1],
    ``cols`` u32[E], rows sorted.  Owns numpy copies of the arrays."""

    def __init__(self, row_ptrs: np.ndarray, cols: np.ndarray):
        self.row_ptrs = np.ascontiguousarray(row_ptrs, dtype=np.uint64)
        self.cols = np.ascontiguousarray(cols, dtype=np.uint32)
        assert int(self.row_ptrs[-1]) == len(self.cols)

    @property
    def nv(self) -> int:
        return len(self.row_ptrs) - 1

    @property
    def ne(self) -> int:
        return len(self.cols)

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptrs)

    # -- handle round trips -------------------------------------------------
    @staticmethod
    def _from_handle(h: C.c_void_p) -> "Graph":
        L = host_lib()
        nv = L.g2h_graph_nv(h)
        ne = L.g2h_graph_ne(h)
        rp = np.ctypeslib.as_array(L.g2h_graph_row_ptrs(h), shape=(nv + 1,)).copy()
        cols = (np.ctypeslib.as_array(L.g2h_graph_cols(h), shape=(ne,)).copy()
                if ne else np.zeros(0, np.uint32))
        L.g2h_graph_free(h)
        return Graph(rp, cols)

    def _handle(self) -> C.c_void_p:
        h = C.c_void_p()
        cols = self.cols if self.ne else np.zeros(1, np.uint32)
        _hcheck(host_lib().g2h_graph_from_csr(self.row_ptrs, self.nv, cols, self.ne, C.byref(h)),
                "g2h_graph_from_csr")
        return h

    @staticmethod
    def read_bin(path: str) -> "Graph":
        h = C.c_void_p()
        _hcheck(host_lib().g2h_graph_read_bin(path.encode(), C.byref(h)), "read_bin")
        return Graph._from_handle(h)

    def write_bin(self, path: str) -> None:
        h = self._handle()
        try:
            _hcheck(host_lib().g2h_graph_write_bin(h, path.encode()), "write_bin")
        finally:
            host_lib().g2h_graph_free(h)

    @staticmethod
    def from_edges(src, dst, nv: int) -> "Graph":
        s = np.ascontiguousarray(src, dtype=np.uint32)
        d = np.ascontiguousarray(dst, dtype=np.uint32)
        if len(s) == 0:
            return Graph(np.zeros(nv + 1, np.uint64), np.zeros(0, np.uint32))
        h = C.c_void_p()
        _hcheck(host_lib().g2h_graph_from_edges(s, d, len(s), nv, C.byref(h)), "from_edges")
        return Graph._from_handle(h)

    def preprocess(self) -> "Graph":
        """Graph::Preprocess (common/graph.h:720-811); returns a new graph."""
        h = self._handle()
        _hcheck(host_lib().g2h_graph_preprocess(h), "preprocess")
        return Graph._from_handle(h)

    def kcore_recode(self):
        """kcore recode (kmeans/kcore.cpp:39-191).  Returns
        (graph, old->new ids, shell offsets, degeneracy)."""
        h = self._handle()
        order = np.zeros(max(self.nv, 1), np.uint32)
        shells = np.zeros(self.nv + 2, np.uint64)
        ns = C.c_uint64()
        dg = C.c_uint32()
        _hcheck(host_lib().g2h_graph_kcore_recode(h, order.ctypes.data, shells.ctypes.data,
                                                  C.byref(ns), C.byref(dg)), "kcore_recode")
        g = Graph._from_handle(h)
        return g, order[: self.nv], shells[: ns.value + 1], dg.value


def kcore_bins(shells: np.ndarray, nb: int, nv: int) -> np.ndarray:
    """Round-robin blocks per k-shell (kmeans/kcore.cpp:193-229)."""
    bins = np.zeros(max(nv, 1), np.uint32)
    sh = np.ascontiguousarray(shells, dtype=np.uint64)
    _hcheck(host_lib().g2h_kcore_bins(sh, len(sh) - 1, nb, bins, nv), "kcore_bins")
    return bins[:nv]


def write_vbmap(path: str, nbins: int, bins: np.ndarray) -> None:
    b = np.ascontiguousarray(bins, dtype=np.uint32)
    _hcheck(host_lib().g2h_write_vbmap(path.encode(), nbins, b, len(b)), "write_vbmap")


def read_vbmap(path: str, nv: int):
    bins = np.zeros(max(nv, 1), np.uint32)
    nb = C.c_uint32()
    _hcheck(host_lib().g2h_read_vbmap(path.encode(), C.byref(nb), bins, nv), "read_vbmap")
    return nb.value, bins[:nv]


def gen_chung_lu(n, gamma, w_min, w_max, target_edges, seed) -> Graph:
    h = C.c_void_p()
    _hcheck(host_lib().g2h_gen_chung_lu(n, gamma, w_min, w_max, target_edges, seed, C.byref(h)),
            "gen_chung_lu")
    return Graph._from_handle(h)


def gen_rmat(scale, edge_factor, a, b, c, seed) -> Graph:
    h = C.c_void_p()
    _hcheck(host_lib().g2h_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)), "gen_rmat")
    return Graph._from_handle(h)


def set_progress(on: bool = True) -> None:
    """progress lines on stderr from long host steps"""
    host_lib().g2h_set_progress(1 if on else 0)


def gen_social(n, gamma, w_min, w_max, backbone_edges, ncomm, smin, smax, comm_alpha,
               pmin, pmax, ecap, seed) -> Graph:
    h = C.c_void_p()
    _hcheck(host_lib().g2h_gen_social(n, gamma, w_min, w_max, backbone_edges, ncomm, smin, smax,
                                      comm_alpha, pmin, pmax, ecap, seed, C.byref(h)), "gen_social")
    return Graph._from_handle(h)


def _qarr(qadj):
    a = (C.c_uint32 * G2_MAXQ)()
    for i, m in enumerate(qadj):
        a[i] = int(m)
    return a


def query_preset(preset: int):
    """Preset pattern (common/graph.h:813-1104) after ReMapVertexId, as
    (n, adjacency bitmasks)."""
    n = C.c_uint32()
    q = (C.c_uint32 * G2_MAXQ)()
    _hcheck(host_lib().g2h_query_preset(preset, C.byref(n), q), "query_preset")
    return n.value, list(q[: n.value])


def query_remap(qadj):
    out = (C.c_uint32 * G2_MAXQ)()
    _hcheck(host_lib().g2h_query_remap(len(qadj), _qarr(qadj), out), "query_remap")
    return list(out[: len(qadj)])


def query_conditions(qadj):
    pairs = (C.c_uint32 * (2 * G2_MAXQ * G2_MAXQ))()
    npairs = C.c_uint32()
    _hcheck(host_lib().g2h_query_conditions(len(qadj), _qarr(qadj), C.byref(npairs), pairs),
            "query_conditions")
    return [(pairs[2 * k], pairs[2 * k + 1]) for k in range(npairs.value)]


def query_automorphisms(qadj) -> int:
    return int(host_lib().g2h_query_automorphisms(len(qadj), _qarr(qadj)))


def query_canonical(qadj) -> int:
    """isomorphism-invariant code of a pattern (equal <=> isomorphic)"""
    return int(host_lib().g2h_query_canonical(len(qadj), _qarr(qadj)))


def query_random(n: int, density: float, seed: int):
    q = (C.c_uint32 * G2_MAXQ)()
    _hcheck(host_lib().g2h_query_random(n, density, seed, q), "query_random")
    return list(q[:n])


def plan_build(qadj) -> G2Plan:
    """Plan::FindRoot .. GenerateStoreStrategy (system/plan.h:51-372) for a
    remapped pattern."""
    p = G2Plan()
    _hcheck(host_lib().g2h_plan_build(len(qadj), _qarr(qadj), C.byref(p)), "plan_build")
    return p


def plan_for_preset(preset: int) -> G2Plan:
    n, q = query_preset(preset)
    return plan_build(q)


def view_materialize(g: Graph, sources: np.ndarray, hop: int, root_degree: int = 0) -> Graph:
    """Hop closure view of a bin (view/view_bin.h:61-159), global ids kept."""
    h = g._handle()
    try:
        rp = np.zeros(g.nv + 1, np.uint64)
        cols = np.zeros(max(g.ne, 1), np.uint32)
        ne = C.c_uint64()
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        if len(src) == 0:
            src = np.zeros(1, np.uint32)
            nsrc = 0
        else:
            nsrc = len(src)
        _hcheck(host_lib().g2h_view_materialize(h, src, nsrc, hop, root_degree, rp, cols,
                                                len(cols), C.byref(ne)), "view_materialize")
    finally:
        host_lib().g2h_graph_free(h)
    return Graph(rp, cols[: ne.value])


from .device import (  # noqa: E402  (device side needs the helpers above)
    Executor,
    hip_lib,
    plan_local_eligible,
)

__all__ = [
    "G2Plan", "Graph", "Executor", "hip_lib", "host_lib", "gen_chung_lu", "gen_rmat",
    "gen_social", "kcore_bins", "write_vbmap", "read_vbmap", "query_preset", "query_remap",
    "query_conditions", "query_automorphisms", "query_random", "plan_build", "plan_for_preset",
    "view_materialize", "plan_local_eligible",
]
