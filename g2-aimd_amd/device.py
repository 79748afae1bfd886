"""Device side of the package: ctypes binding of libg2hip.so (include/g2hip.h).

``Executor`` mirrors the reference's per-GPU PipelineExecutor
(system/pipeline_executor.h:44-207): ``transfer`` uploads a view (CSR rows by
global vertex id + source vertices, ``Transfer`` :121-149), ``run_mce`` is
``Run`` with the BK app (app_BK/BK.h), ``set_plan`` + ``run_gm`` are ``Run``
with the GMatch app (app_gmatch/gmatch.h).  Errors raise RuntimeError with the
library's message (the reference prints and exits, device/cuda_context.h:18-25).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_hip = None

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

G2_RUN_ACCOUNT = 1
G2_RUN_CANONICAL = 2
G2_RUN_LOCAL_BYTES = 4

# every symbol include/g2hip.h declares
EXPORTS = ("g2_create", "g2_set_tuning", "g2_destroy", "g2_last_error", "g2_upload_view", "g2_mce_run",
           "g2_gm_set_plan", "g2_gm_run", "g2_device_info", "g2_device_count", "g2_pinned_alloc",
           "g2_pinned_free", "g2_plan_local_eligible", "g2_features", "g2_mlp_create",
           "g2_mlp_destroy", "g2_mlp_get_params", "g2_mlp_set_params", "g2_mlp_infer", "g2_mlp_grad",
           "g2_mlp_train", "g2_gm_run_batch", "g2_gm_autotune", "g2_core_numbers", "g2_root_inputs",
           "g2_gm_run_noalloc", "g2_noalloc_last_error", "g2_learned_schedule")
G2_NFEAT_INT = 5
G2_MLP_NPARAM = 4801


class Tuning(C.Structure):
    _fields_ = [("waves_per_cu", C.c_uint32), ("block_waves", C.c_uint32),
                ("wave_scratch_words", C.c_uint64), ("big_waves", C.c_uint32),
                ("gm_chunk", C.c_uint32), ("mce_small_max", C.c_uint32),
                ("mce_donate_every", C.c_uint32), ("mce_roots_per_wave", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("launches", C.c_uint32), ("total_ms", C.c_double),
                ("nodes", C.c_uint64), ("alg_bytes", C.c_uint64), ("big_roots", C.c_uint64),
                ("max_scratch_words", C.c_uint64), ("wave_idle_frac", C.c_double),
                ("donations", C.c_uint64), ("local_bitmap", C.c_uint32), ("reserved", C.c_uint32),
                ("local_pass_roots", C.c_uint64 * 5), ("host_spill_bytes", C.c_uint64),
                ("device_bytes", C.c_uint64), ("local_pass_ms", C.c_double * 5),
                ("local_pass_bytes", C.c_uint64 * 5)]

    def to_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}
        for k in ("local_pass_roots", "local_pass_ms", "local_pass_bytes"):
            d[k] = list(getattr(self, k))
        return d


class Sched(C.Structure):
    _fields_ = [("order", C.c_void_p), ("n", C.c_uint64), ("begin", C.c_uint64),
                ("end", C.c_uint64), ("split", C.c_void_p)]


class Batch(C.Structure):
    _fields_ = [("order", C.c_void_p), ("roots_per_task", C.c_uint32), ("window", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("reserved", C.c_uint32), ("node_budget", C.c_uint64)]

    def to_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k not in ("order", "reserved")}


def hip_lib() -> C.CDLL:
    """Load libg2hip.so (raises if it was not built: there is no fallback)."""
    global _hip
    if _hip is not None:
        return _hip
    path = os.path.join(_HERE, "libg2hip.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    from . import G2Plan  # noqa: circular-safe (package already initialised)

    vp = C.c_void_p
    L.g2_create.argtypes = [C.POINTER(vp), C.c_int, C.c_uint64, C.POINTER(Tuning)]
    L.g2_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.g2_destroy.argtypes = [vp]
    L.g2_destroy.restype = None
    L.g2_last_error.argtypes = [vp]
    L.g2_last_error.restype = C.c_char_p
    L.g2_upload_view.argtypes = [vp, u64p, C.c_uint64, u32p, C.c_uint64, u32p, C.c_uint64]
    L.g2_mce_run.argtypes = [vp, C.POINTER(Sched), C.c_uint32, C.POINTER(C.c_uint64), vp,
                             C.c_uint32, vp, C.POINTER(Stats)]
    L.g2_gm_set_plan.argtypes = [vp, C.POINTER(G2Plan)]
    L.g2_gm_run.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint64), vp, C.POINTER(Stats)]
    L.g2_device_info.argtypes = [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                 C.c_char_p, C.c_uint32]
    L.g2_plan_local_eligible.argtypes = [C.POINTER(G2Plan)]
    L.g2_gm_run_batch.argtypes = [vp, C.POINTER(G2Plan), C.c_uint32, C.POINTER(Batch), u64p, u64p,
                                  u32p, C.POINTER(Stats)]
    L.g2_gm_autotune.argtypes = [vp, C.POINTER(G2Plan), C.c_uint32, C.c_uint32, C.c_uint64,
                                 C.POINTER(Batch), C.POINTER(C.c_double)]
    f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.g2_features.argtypes = [vp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
    L.g2_core_numbers.argtypes = [vp, u32p]
    L.g2_root_inputs.argtypes = [vp, vp, f32p]
    L.g2_mlp_create.argtypes = [C.POINTER(vp), C.c_int, C.c_uint64]
    L.g2_mlp_destroy.argtypes = [vp]
    L.g2_mlp_destroy.restype = None
    L.g2_mlp_get_params.argtypes = [vp, f32p]
    L.g2_mlp_set_params.argtypes = [vp, f32p]
    L.g2_mlp_infer.argtypes = [vp, f32p, C.c_uint64, f32p]
    L.g2_mlp_grad.argtypes = [vp, f32p, f32p, C.c_uint64, f32p, C.POINTER(C.c_float)]
    L.g2_mlp_train.argtypes = [vp, f32p, f32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                               C.POINTER(C.c_float)]
    L.g2_learned_schedule.argtypes = [vp, vp, f32p, f32p, vp, C.c_uint64, C.c_uint32, u32p, u32p, vp]
    _hip = L
    return L


def plan_local_eligible(plan) -> bool:
    """Host-only check: is the plan counted by the oriented local-bitmap kernel?"""
    rc = hip_lib().g2_plan_local_eligible(C.byref(plan))
    if rc < 0:
        raise RuntimeError("invalid plan")
    return rc == 1


def device_info(device: int = 0) -> dict:
    cus, mem = C.c_uint32(), C.c_uint64()
    name = C.create_string_buffer(64)
    rc = hip_lib().g2_device_info(device, C.byref(cus), C.byref(mem), name, 64)
    if rc != 0:
        raise RuntimeError(f"g2_device_info failed ({rc})")
    return {"num_cus": cus.value, "total_mem": mem.value, "arch": name.value.decode()}


class Executor:
    """One device context (a consumer of the reference's worker pipeline)."""

    def __init__(self, device: int = 0, mem_budget: int = 0, **tuning):
        L = hip_lib()
        t = Tuning(**tuning)
        self._ctx = C.c_void_p()
        rc = L.g2_create(C.byref(self._ctx), device, mem_budget, C.byref(t))
        if rc != 0:
            raise RuntimeError(f"g2_create failed ({rc}): no usable GPU or bad arguments")
        self.ns = 0
        self.nv = 0

    def set_tuning(self, **tuning) -> None:
        """New launch tuning for the next runs (the view stays uploaded)."""
        t = Tuning(**tuning)
        self._check(hip_lib().g2_set_tuning(self._ctx, C.byref(t)), "g2_set_tuning")

    def close(self) -> None:
        if self._ctx:
            hip_lib().g2_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = hip_lib().g2_last_error(self._ctx).decode()
            raise RuntimeError(f"{what} failed ({rc}): {msg}")

    def transfer(self, g, sources=None) -> None:
        """Upload a view (``Transfer``): CSR rows by global id + source ids."""
        src = (np.arange(g.nv, dtype=np.uint32) if sources is None
               else np.ascontiguousarray(sources, dtype=np.uint32))
        cols = g.cols if g.ne else np.zeros(1, np.uint32)
        s = src if len(src) else np.zeros(1, np.uint32)
        self._check(hip_lib().g2_upload_view(self._ctx, g.row_ptrs, g.nv, cols, g.ne, s, len(src)),
                    "g2_upload_view")
        self.ns = len(src)
        self.nv = g.nv

    def run_mce(self, hist_len: int = 256, per_root: bool = False, order=None, split=None,
                account: bool = False, begin: int = 0, end: int = 0) -> dict:
        """Maximal clique enumeration over the uploaded sources.  `order`
        (permutation of source positions) and `split` (per source: parts its
        first-level children are dealt into) only change the schedule."""
        total = C.c_uint64()
        hist = np.zeros(hist_len, np.uint64)
        prn = np.zeros(max(self.ns, 1), np.uint64) if per_root else None
        st = Stats()
        sched = None
        if order is not None or split is not None or end > begin:
            o = sp = None
            if order is not None:
                o = np.ascontiguousarray(order, dtype=np.uint32)
                if len(o) != self.ns:
                    raise ValueError("order must be a permutation of the sources")
            if split is not None:
                sp = np.ascontiguousarray(split, dtype=np.uint32)
                if len(sp) != self.ns:
                    raise ValueError("split needs one entry per source")
            sched = Sched(o.ctypes.data if o is not None else None, self.ns, begin, end,
                          sp.ctypes.data if sp is not None else None)
            self._sched_keep = (o, sp)
        rc = hip_lib().g2_mce_run(self._ctx, C.byref(sched) if sched else None,
                                  G2_RUN_ACCOUNT if account else 0, C.byref(total),
                                  hist.ctypes.data, hist_len,
                                  prn.ctypes.data if per_root else None, C.byref(st))
        self._check(rc, "g2_mce_run")
        out = {"total": total.value, "hist": hist, "stats": st.to_dict()}
        if per_root:
            out["per_root_nodes"] = prn[: self.ns]
        return out

    def set_plan(self, plan) -> None:
        self._check(hip_lib().g2_gm_set_plan(self._ctx, C.byref(plan)), "g2_gm_set_plan")

    def run_gm(self, per_root: bool = False, account: bool = False, canonical: bool = False,
               local_bytes: bool = False) -> dict:
        """Match count; whole-graph totals of clique-like plans use the oriented
        local-bitmap kernel unless `canonical` (or per-root / byte accounting).
        `local_bytes`: the local path also reports its per-pass algorithmic bytes."""
        cnt = C.c_uint64()
        prc = np.zeros(max(self.ns, 1), np.uint64) if per_root else None
        st = Stats()
        flags = ((G2_RUN_ACCOUNT if account else 0) | (G2_RUN_CANONICAL if canonical else 0) |
                 (G2_RUN_LOCAL_BYTES if local_bytes else 0))
        rc = hip_lib().g2_gm_run(self._ctx, flags, C.byref(cnt),
                                 prc.ctypes.data if per_root else None, C.byref(st))
        self._check(rc, "g2_gm_run")
        out = {"count": cnt.value, "stats": st.to_dict()}
        if per_root:
            out["per_root_count"] = prc[: self.ns]
        return out

    @staticmethod
    def _plans(plans):
        from . import G2Plan
        arr = (G2Plan * max(len(plans), 1))()
        for i, p in enumerate(plans):
            arr[i] = p
        return arr

    def run_gm_batch(self, plans, order=None, roots_per_task: int = 1, window: int = 0,
                     waves_per_cu: int = 0, node_budget: int = 0) -> dict:
        """Counts of many plans in one persistent launch (canonical traversal):
        counts[q] equals run_gm(canonical=True) of plans[q] alone."""
        nq = len(plans)
        arr = self._plans(plans)
        o = None
        if order is not None:
            o = np.ascontiguousarray(order, dtype=np.uint32)
            if len(o) != nq:
                raise ValueError("order must be a permutation of the queries")
        b = Batch(o.ctypes.data if o is not None else None, roots_per_task, window, waves_per_cu, 0,
                  node_budget)
        counts = np.zeros(max(nq, 1), np.uint64)
        nodes = np.zeros(max(nq, 1), np.uint64)
        over = np.zeros(max(nq, 1), np.uint32)
        st = Stats()
        self._check(hip_lib().g2_gm_run_batch(self._ctx, arr, nq, C.byref(b), counts, nodes, over,
                                              C.byref(st)), "g2_gm_run_batch")
        return {"counts": counts[:nq], "nodes": nodes[:nq], "over_budget": over[:nq].astype(bool),
                "stats": st.to_dict()}

    def autotune(self, plans, sample_queries: int = 0, seed: int = 1, node_budget: int = 0) -> dict:
        """Launch shape of run_gm_batch picked by timing a query sample."""
        arr = self._plans(plans)
        b = Batch(None, 0, 0, 0, 0, node_budget)
        ms = C.c_double()
        self._check(hip_lib().g2_gm_autotune(self._ctx, arr, len(plans), sample_queries, seed,
                                             C.byref(b), C.byref(ms)), "g2_gm_autotune")
        out = b.to_dict()
        out["sample_ms"] = ms.value
        return out

    def core_numbers(self) -> np.ndarray:
        """core number of every vertex of the uploaded view (device h-index iteration)"""
        out = np.zeros(max(self.nv, 1), np.uint32)
        self._check(hip_lib().g2_core_numbers(self._ctx, out), "g2_core_numbers")
        return out[: self.nv]

    def root_inputs(self, core=None) -> np.ndarray:
        """the cost model's 8 fp32 inputs of every source, computed on the device"""
        out = np.zeros((max(self.ns, 1), 8), np.float32)
        cr = None if core is None else np.ascontiguousarray(core, dtype=np.uint32)
        self._check(hip_lib().g2_root_inputs(self._ctx, cr.ctypes.data if cr is not None else None, out),
                    "g2_root_inputs")
        return out[: self.ns]

    def learned_schedule(self, mlp, mu, sd, waves: int, core=None, max_split: int = 0):
        """(order, split, pred) of the uploaded sources from the cost model,
        computed on the device (g2_learned_schedule)."""
        n = max(self.ns, 1)
        order = np.zeros(n, np.uint32)
        split = np.zeros(n, np.uint32)
        pred = np.zeros(n, np.float32)
        cr = None if core is None else np.ascontiguousarray(core, dtype=np.uint32)
        self._check(hip_lib().g2_learned_schedule(
            self._ctx, mlp._h, np.ascontiguousarray(mu, dtype=np.float32), np.ascontiguousarray(sd, dtype=np.float32),
            cr.ctypes.data if cr is not None else None, waves, max_split, order, split, pred.ctypes.data),
            "g2_learned_schedule")
        return order[: self.ns], split[: self.ns], pred[: self.ns]

    def features(self) -> np.ndarray:
        """Integer root features (ns x 5): degree, distinct 2-hop ball size
        (HyperLogLog estimate), triangles, neighbours sharing a neighbour with
        the root, later neighbours."""
        out = np.zeros((max(self.ns, 1), G2_NFEAT_INT), np.int64)
        self._check(hip_lib().g2_features(self._ctx, out), "g2_features")
        return out[: self.ns]


class Mlp:
    """Cost model 8 -> 64 -> 64 -> 1 trained and run by the matrix-core kernels."""

    def __init__(self, device: int = 0, seed: int = 1):
        self._h = C.c_void_p()
        if hip_lib().g2_mlp_create(C.byref(self._h), device, seed) != 0:
            raise RuntimeError("g2_mlp_create failed")

    def close(self):
        if self._h:
            hip_lib().g2_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _x(x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.ndim == 2 and x.shape[1] == 8
        return x

    def params(self) -> np.ndarray:
        p = np.zeros(G2_MLP_NPARAM, np.float32)
        assert hip_lib().g2_mlp_get_params(self._h, p) == 0
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert len(p) == G2_MLP_NPARAM and hip_lib().g2_mlp_set_params(self._h, p) == 0

    def infer(self, x) -> np.ndarray:
        x = self._x(x)
        y = np.zeros(max(len(x), 1), np.float32)
        if len(x):
            assert hip_lib().g2_mlp_infer(self._h, x, len(x), y) == 0
        return y[: len(x)]

    def grad(self, x, y):
        x = self._x(x)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(G2_MLP_NPARAM, np.float32)
        loss = C.c_float()
        assert hip_lib().g2_mlp_grad(self._h, x, y, len(x), g, C.byref(loss)) == 0
        return loss.value, g

    def train(self, x, y, epochs=20, batch=1024, lr=1e-3) -> float:
        x = self._x(x)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_float()
        rc = hip_lib().g2_mlp_train(self._h, x, y, len(x), epochs, batch, lr, C.byref(loss))
        if rc != 0:
            raise RuntimeError(f"g2_mlp_train failed ({rc})")
        return loss.value

