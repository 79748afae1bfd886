"""Learned root scheduling for maximal clique enumeration.

The reference runs roots in source order, one warp per queued subgraph
(app_BK/BK.h:323-355, system/pipeline_executor.h:151-207).  Here a root's
whole subtree is one wave's work, so a few heavy roots can leave most waves
idle at the end of a launch.  The scheduler
  1. computes 8 features per root (GPU feature kernel + kcore shells),
  2. predicts log(1 + subtree nodes) with the MFMA MLP cost model, trained on
     the per-root node counts the plain enumerator logs,
  3. orders roots by predicted cost (longest first) and deals the first-level
     children of predicted-heavy roots over several waves.
Neither step changes a result: the kernel's tree, counts, histogram and
per-root node totals are the same for every order and split (tests).  The
plain source order with no split is the fallback and the yardstick.

The feature assembly, standardisation, schedule rule, core numbers and model
files are C++ in libg2host.so (host/sched.cpp, include/g2host.h), shared with
the `bk` command line; this module binds them for the tests and bench.py.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import host_lib

NFEAT = 8
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class Norm(C.Structure):
    """g2h_norm: per-feature mean and scale"""
    _fields_ = [("mu", C.c_float * NFEAT), ("sd", C.c_float * NFEAT)]


def _lib():
    L = host_lib()
    if not getattr(L, "_sched_bound", False):
        L.g2h_core_numbers.argtypes = [_u64p, C.c_uint64, _u32p, _u32p]
        L.g2h_root_features.argtypes = [_i64p, _u32p, C.c_uint64, _u32p, C.c_uint64, _f32p]
        L.g2h_norm_fit.argtypes = [_f32p, C.c_uint64, C.POINTER(Norm)]
        L.g2h_norm_apply.argtypes = [C.POINTER(Norm), _f32p, C.c_uint64]
        L.g2h_norm_apply.restype = None
        L.g2h_learned_schedule.argtypes = [_f32p, _i64p, C.c_uint64, C.c_uint64, C.c_uint32, _u32p, _u32p]
        L.g2h_model_save.argtypes = [C.c_char_p, C.POINTER(Norm), _f32p, C.c_uint64]
        L.g2h_model_load.argtypes = [C.c_char_p, C.POINTER(Norm), _f32p, C.c_uint64]
        L._sched_bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {host_lib().g2h_last_error().decode()}")

FEATURE_NAMES = ("log1p_degree", "log1p_core", "log1p_2hop_ball", "log1p_triangles", "clustering",
                 "degeneracy_rank", "shared_neighbour_fraction", "log1p_later_neighbours")


def core_numbers(shells: np.ndarray, nv: int) -> np.ndarray:
    """core number of every vertex of a kcore-recoded graph (shell index)"""
    core = np.zeros(nv, np.int64)
    for k in range(len(shells) - 1):
        core[int(shells[k]):int(shells[k + 1])] = k
    return core


def core_numbers_csr(g) -> np.ndarray:
    """core number of every vertex of any CSR graph (bucket peel, C++)"""
    out = np.zeros(max(g.nv, 1), np.uint32)
    cols = g.cols if g.ne else np.zeros(1, np.uint32)
    _check(_lib().g2h_core_numbers(g.row_ptrs, g.nv, cols, out), "g2h_core_numbers")
    return out[: g.nv]


def root_features(ex, sources: np.ndarray, core: np.ndarray, nv: int) -> np.ndarray:
    """float32 (ns, 8) model inputs of the uploaded sources"""
    f = np.ascontiguousarray(ex.features(), dtype=np.int64)
    src = np.ascontiguousarray(sources, dtype=np.uint32)
    cr = np.ascontiguousarray(core, dtype=np.uint32)
    out = np.zeros((max(len(src), 1), NFEAT), np.float32)
    if len(src):
        _check(_lib().g2h_root_features(f, src, len(src), cr, nv, out), "g2h_root_features")
    return out[: len(src)]


def learned_schedule(pred_log_nodes: np.ndarray, later_neighbours: np.ndarray, waves: int,
                     max_split: int = 4096):
    """(order, split) from predicted log(1+nodes): longest predicted first;
    a root predicted to exceed the per-wave share is dealt into
    ceil(cost / share) parts (at most its number of later neighbours, the
    most first-level children it can have)."""
    pred = np.ascontiguousarray(pred_log_nodes, dtype=np.float32)
    later = np.ascontiguousarray(later_neighbours, dtype=np.int64)
    n = len(pred)
    order = np.zeros(max(n, 1), np.uint32)
    split = np.zeros(max(n, 1), np.uint32)
    _check(_lib().g2h_learned_schedule(pred if n else np.zeros(1, np.float32),
                                       later if n else np.zeros(1, np.int64), n, waves, max_split,
                                       order, split), "g2h_learned_schedule")
    return order[:n], split[:n]


def spearman(a: np.ndarray, b: np.ndarray) -> float:
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    den = np.sqrt((ra * ra).sum() * (rb * rb).sum())
    return float((ra * rb).sum() / den) if den > 0 else 0.0


class Standardizer:
    """per-feature affine map fitted on the training rows (g2h_norm_fit)"""

    def __init__(self, x=None, norm: Norm | None = None):
        self.norm = Norm() if norm is None else norm
        if x is not None:
            xx = np.ascontiguousarray(x, dtype=np.float32)
            _check(_lib().g2h_norm_fit(xx, len(xx), C.byref(self.norm)), "g2h_norm_fit")

    @property
    def mu(self):
        return np.array(self.norm.mu[:], np.float32)

    @property
    def sd(self):
        return np.array(self.norm.sd[:], np.float32)

    def __call__(self, x: np.ndarray) -> np.ndarray:
        out = np.array(x, dtype=np.float32, order="C", copy=True)
        _lib().g2h_norm_apply(C.byref(self.norm), out, len(out))
        return out


def save_model(path: str, norm: Standardizer, params: np.ndarray) -> None:
    p = np.ascontiguousarray(params, dtype=np.float32)
    _check(_lib().g2h_model_save(path.encode(), C.byref(norm.norm), p, len(p)), "g2h_model_save")


def load_model(path: str, nparams: int):
    """(Standardizer, parameters) of a model file written by save_model / bk"""
    n = Norm()
    p = np.zeros(nparams, np.float32)
    _check(_lib().g2h_model_load(path.encode(), C.byref(n), p, nparams), "g2h_model_load")
    return Standardizer(norm=n), p


def simulated_makespan(cost: np.ndarray, order: np.ndarray, split: np.ndarray, waves: int) -> float:
    """greedy list schedule of per-task costs on `waves` identical workers in
    the kernel's fetch order (a split root's cost is dealt evenly over its
    parts): the makespan the dynamic fetch would reach if each wave's time
    were proportional to its nodes"""
    import heapq
    load = [0.0] * waves
    heapq.heapify(load)
    for i in order:
        k = int(split[i]) if split is not None else 1
        part = float(cost[i]) / k
        for _ in range(k):
            t = heapq.heappop(load)
            heapq.heappush(load, t + part)
    return max(load)


def prepare_graph(g2, scale: int, seed: int, edge_factor: int = 16):
    """R-MAT graph of the matching workload family, degeneracy-ordered"""
    g = g2.gen_rmat(scale, edge_factor, 0.57, 0.19, 0.19, seed).preprocess()
    g, _, shells, _ = g.kcore_recode()
    return g, core_numbers(shells, g.nv)


def logged_rows(dev, g, core, **tuning):
    """features and per-root log(1 + subtree nodes) logged by a plain run"""
    with dev.Executor(device=0, **tuning) as ex:
        ex.transfer(g)
        x = root_features(ex, np.arange(g.nv, dtype=np.uint32), core, g.nv)
        r = ex.run_mce(per_root=True)
    return x, np.log1p(r["per_root_nodes"].astype(np.float64)).astype(np.float32), r


def experiment(g2, dev, scale: int, train_seeds, test_seed: int, epochs: int = 30, batch: int = 512,
               lr: float = 2e-3, reps: int = 3, log=print, **tuning) -> dict:
    """Train the cost model on plain-run logs of the training seeds, schedule
    the test seed's roots with it and compare against the plain order."""
    xs, ys = [], []
    for s in train_seeds:
        g, core = prepare_graph(g2, scale, s)
        x, y, _ = logged_rows(dev, g, core, **tuning)
        xs.append(x)
        ys.append(y)
        log(f"train seed {s:#x}: {g.nv} roots, max log1p(nodes) {y.max():.2f}")
    X, Y = np.concatenate(xs), np.concatenate(ys)
    norm = Standardizer(X)
    mlp = dev.Mlp(device=0, seed=1)
    loss = mlp.train(norm(X), Y, epochs=epochs, batch=batch, lr=lr)
    log(f"trained on {len(X)} roots: last-epoch mse {loss:.4f}")

    g, core = prepare_graph(g2, scale, test_seed)
    out = {"scale": scale, "train_roots": int(len(X)), "test_roots": int(g.nv), "train_mse": loss}
    with dev.Executor(device=0, **tuning) as ex:
        ex.transfer(g)
        x = root_features(ex, np.arange(g.nv, dtype=np.uint32), core, g.nv)
        pred = mlp.infer(norm(x))
        plain = ex.run_mce(per_root=True)
        truth = np.log1p(plain["per_root_nodes"].astype(np.float64))
        info = dev.device_info(0)
        waves = info["num_cus"] * max(1, tuning.get("waves_per_cu", 0) or 16)
        later = ex.features()[:, 4]
        order, split = learned_schedule(pred, later, waves)
    mlp.close()
    nodes = plain["per_root_nodes"].astype(np.float64)
    out.update({
        "max_cliques": int(plain["total"]),
        "rmse_log_nodes": float(np.sqrt(np.mean((pred - truth) ** 2))),
        "spearman": spearman(pred, truth),
        "split_roots": int((split > 1).sum()), "split_parts": int(split.sum()),
        "sim_makespan_plain": simulated_makespan(nodes, np.arange(g.nv), None, waves),
        "sim_makespan_learned": simulated_makespan(nodes, order, split, waves),
        "sim_makespan_bound": max(nodes.sum() / waves, 1.0),
    })
    # wall time of the plain and learned schedules, with and without work
    # donation.  With donation the learned form keeps every root whole (the
    # order only): donation divides a heavy root where its work actually is,
    # while first-level parts each rebuild the root's bitsets
    whole = np.ones_like(split)
    for dname, extra, lsplit in (("donate", {}, whole), ("nodonate", {"mce_donate_every": 0xFFFFFFFF}, split)):
        with dev.Executor(device=0, **{**tuning, **extra}) as ex:
            ex.transfer(g)
            for sname, kw in (("plain", {}), ("learned", {"order": order, "split": lsplit})):
                r = ex.run_mce(**kw)
                if r["total"] != plain["total"] or not np.array_equal(r["hist"], plain["hist"]):
                    raise RuntimeError("a schedule changed the result")
                st = [ex.run_mce(**kw)["stats"] for _ in range(reps)]
                key = f"{sname}_{dname}"
                out[key + "_ms"] = float(np.median([s["kernel_ms"] for s in st]))
                out[key + "_idle_frac"] = float(np.median([s["wave_idle_frac"] for s in st]))
                out[key + "_donations"] = int(st[-1]["donations"])
                log(f"{key}: {out[key + '_ms']:.2f} ms, idle {out[key + '_idle_frac']:.3f}")
    out["speedup_learned_vs_plain"] = out["plain_donate_ms"] / out["learned_donate_ms"]
    out["speedup_learned_vs_plain_nodonate"] = out["plain_nodonate_ms"] / out["learned_nodonate_ms"]
    return out


MCE_SHAPES = ({"mce_roots_per_wave": 1, "waves_per_cu": 16}, {"mce_roots_per_wave": 2, "waves_per_cu": 16},
              {"mce_roots_per_wave": 2, "waves_per_cu": 24}, {"mce_roots_per_wave": 4, "waves_per_cu": 24},
              {"mce_roots_per_wave": 4, "waves_per_cu": 32})


def autotune_mce(ex, shapes=MCE_SHAPES, reps: int = 1, log=print, **run_kw) -> tuple:
    """Launch shape (roots per wave, resident waves) of the maximal-clique
    run with the lowest kernel time on the uploaded view; every trial's
    result is checked equal.  Returns (shape, ms, trials)."""
    best, best_ms, ref, trials = None, None, None, []
    for shape in shapes:
        ex.set_tuning(**shape)
        ms = []
        for _ in range(reps + 1):
            r = ex.run_mce(**run_kw)
            ms.append(r["stats"]["kernel_ms"])
        key = (r["total"], r["hist"].tobytes())
        if ref is None:
            ref = key
        elif key != ref:
            raise RuntimeError("a launch shape changed the result")
        t = float(np.min(ms[1:]))
        trials.append({**shape, "ms": t})
        log(f"shape {shape}: {t:.2f} ms")
        if best_ms is None or t < best_ms:
            best, best_ms = shape, t
    ex.set_tuning(**best)
    return best, best_ms, trials
