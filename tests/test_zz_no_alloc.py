"""g2_gm_run_noalloc (the allocation-free matching entry point) against
g2_gm_run's canonical traversal over the same view: two calls in a row, the
second on a smaller input, and the errors it reports instead of running."""
import ctypes as C

import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu
RMAT_SEED = 0x6732A1D0C0000002


def bind(g2):
    L = g2.hip_lib()
    L.g2_gm_run_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                    C.c_void_p, C.c_uint64, C.POINTER(g2.G2Plan), C.POINTER(C.c_uint64)]
    L.g2_noalloc_last_error.restype = C.c_char_p
    return L


class DeviceArrays:
    """The view's arrays in device memory, allocated and filled through the
    HIP runtime libg2hip.so itself loaded (the test's plumbing; the entry
    point under test allocates nothing)."""

    def __init__(self, g2, arrays):
        self.hip = hip_runtime(g2)
        self.ptrs = []
        for a in arrays:
            a = np.ascontiguousarray(a)
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
            if a.nbytes:
                assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
            self.ptrs.append(p)

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def hip_runtime(g2):
    """the libamdhip64 that libg2hip.so loaded (one HIP runtime per process)"""
    g2.hip_lib()
    for line in open("/proc/self/maps"):
        path = line.split()[-1]
        if "libamdhip64.so" in path:
            hip = C.CDLL(path)
            hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            hip.hipFree.argtypes = [C.c_void_p]
            return hip
    raise RuntimeError("libg2hip.so did not load a HIP runtime")


def noalloc_raw(L, g2, row_ptrs, cols, sources, plan):
    nv = len(row_ptrs) - 1
    d = DeviceArrays(g2, [row_ptrs, cols if len(cols) else np.zeros(1, np.uint32),
                          np.ascontiguousarray(sources, np.uint32)])
    out = C.c_uint64()
    try:
        rc = L.g2_gm_run_noalloc(0, d.ptrs[0], nv, d.ptrs[1], len(cols), d.ptrs[2], len(sources),
                                 C.byref(plan), C.byref(out))
    finally:
        d.free()
    return rc, out.value, L.g2_noalloc_last_error().decode()


def noalloc(L, g2, g, sources, plan):
    return noalloc_raw(L, g2, g.row_ptrs, g.cols, sources, plan)


def canonical(g2, g, sources, plan):
    with g2.Executor(device=0) as ex:
        ex.transfer(g, sources)
        ex.set_plan(plan)
        return ex.run_gm(canonical=True)["count"]


def test_noalloc_equals_run(g2):
    print("case: noalloc vs g2_gm_run, two calls", flush=True)
    L = bind(g2)
    # (a sparse random pattern on a hub-heavy graph has ~1e11 matches at
    # R-MAT 13: the pattern runs on the smaller graphs only)
    cases = [(g2.plan_for_preset(24), g2.gen_rmat(13, 16, 0.57, 0.19, 0.19, RMAT_SEED).preprocess(), 1),
             (g2.plan_for_preset(24), g2.gen_rmat(9, 8, 0.57, 0.19, 0.19, RMAT_SEED).preprocess(), 2),
             (g2.plan_build(g2.query_remap(g2.query_random(5, 0.6, 40007))), G.gnp(g2, 300, 0.05, 8), 1),
             (g2.plan_build(g2.query_remap(g2.query_random(5, 0.6, 40007))),
              g2.gen_rmat(9, 8, 0.57, 0.19, 0.19, RMAT_SEED).preprocess(), 3)]
    for plan, g, step in cases:
        for s in (np.arange(0, g.nv, step, dtype=np.uint32),):
            rc, got, msg = noalloc(L, g2, g, s, plan)
            assert rc == 0, msg
            ref = canonical(g2, g, s, plan)
            print(f"nv={g.nv} ns={len(s)} n={plan.n}: {got} (g2_gm_run {ref})", flush=True)
            assert got == ref


def test_noalloc_errors(g2):
    print("case: noalloc errors", flush=True)
    L = bind(g2)
    plan = g2.plan_for_preset(24)
    g = G.complete(g2, 10)
    src = np.arange(10, dtype=np.uint32)
    # a column id beyond the vertex count, and a decreasing row_ptrs, are found before the search
    cols = g.cols.copy()
    cols[5] = 12
    rc, _, msg = noalloc_raw(L, g2, g.row_ptrs, cols, src, plan)
    assert rc == -1 and "column id out of range" in msg, msg
    rp = g.row_ptrs.copy()
    rp[3], rp[4] = rp[4], rp[3]
    rc, _, msg = noalloc_raw(L, g2, rp, g.cols, src, plan)
    assert rc == -1 and "row_ptrs" in msg, msg
    # a star whose hub row needs a wave arena beyond the 64 MiB array
    leaves = 2_200_000
    hub = np.zeros(leaves, np.uint32)
    star = g2.Graph.from_edges(hub, np.arange(1, leaves + 1, dtype=np.uint32), leaves + 1)
    rc, _, msg = noalloc(L, g2, star, np.arange(4, dtype=np.uint32), plan)
    assert rc == -3 and "64 MiB" in msg
    # after the errors the array serves a valid call again
    rc, got, msg = noalloc(L, g2, g, src, plan)
    assert rc == 0 and got == 210, msg
