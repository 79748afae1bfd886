// csrc/g2_ctx.hip -- device context, memory budget and the C ABI entry points.
//
// Replaces the reference executor's setup (system/pipeline_executor.h:44-99,
// system/appbase.h:93-114), which grabs all free device memory for its level
// buffers: here every allocation is charged to an explicit per-context budget
// so that several consumers can share one MI355X, and buffers grow on demand.
#include <algorithm>
#include <cstring>

#include "g2_internal.hpp"

namespace g2 {

int set_err(g2_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_err(g2_ctx* c, hipError_t e, const char* what) {
    return set_err(c, G2_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int ensure(g2_ctx* c, DevBuf& b, uint64_t bytes, const char* what) {
    if (bytes <= b.bytes) return G2_OK;
    bytes = (bytes + 255) & ~255ull;
    const uint64_t after = c->used - b.bytes + bytes;
    if (c->budget && after > c->budget)
        return set_err(c, G2_EBUDGET, std::string(what) + ": needs " + std::to_string(bytes) +
                                          " B, budget left " + std::to_string(c->budget - (c->used - b.bytes)));
    if (b.ptr) {
        hipError_t e = hipFree(b.ptr);
        if (e != hipSuccess) return hip_err(c, e, "hipFree");
        c->used -= b.bytes;
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) return hip_err(c, e, what);
    b.bytes = bytes;
    c->used += bytes;
    return G2_OK;
}

static void release(g2_ctx* c, DevBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    c->used -= b.bytes;
    b.ptr = nullptr;
    b.bytes = 0;
}

static void apply_tuning_defaults(g2_ctx* c) {
    if (!c->tune.block_waves) c->tune.block_waves = 4;
    if (!c->tune.wave_scratch_words) c->tune.wave_scratch_words = 1ull << 16;
    if (!c->tune.big_waves) c->tune.big_waves = 256;
    if (!c->tune.mce_small_max) c->tune.mce_small_max = 64;
    if (!c->tune.mce_donate_every) c->tune.mce_donate_every = 256;
}

}  // namespace g2

using namespace g2;

extern "C" {

int g2_device_info(int device, uint32_t* num_cus, uint64_t* total_mem, char* name, uint32_t name_len) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return G2_EHIP;
    if (num_cus) *num_cus = p.multiProcessorCount;
    if (total_mem) *total_mem = p.totalGlobalMem;
    if (name && name_len) {
        std::strncpy(name, p.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    return G2_OK;
}

int g2_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void* g2_pinned_alloc(uint64_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void g2_pinned_free(void* p) {
    if (p) (void)hipHostFree(p);
}

int g2_create(g2_ctx** out, int device, uint64_t budget, const g2_tuning* t) {
    if (!out) return G2_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return G2_EHIP;
    if (device < 0 || device >= ndev) return G2_EINVAL;
    auto* c = new g2_ctx;
    c->device = device;
    c->budget = budget;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete c; return G2_EHIP; }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) { delete c; return G2_EHIP; }
    c->num_cus = p.multiProcessorCount;
    if (t) c->tune = *t;
    apply_tuning_defaults(c);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return G2_EHIP; }
    for (auto& ev : c->ev)
        if (hipEventCreate(&ev) != hipSuccess) { g2_destroy(c); return G2_EHIP; }
    *out = c;
    return G2_OK;
}

int g2_set_tuning(g2_ctx* c, const g2_tuning* t) {
    if (!c || !t) return G2_EINVAL;
    const uint32_t chunk = c->tune.gm_chunk;
    c->tune = *t;
    c->tune.gm_chunk = chunk;  // matching windows are fixed at upload
    apply_tuning_defaults(c);
    return G2_OK;
}

void g2_destroy(g2_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : {&c->rp, &c->cols, &c->src, &c->gm_task_off, &c->counters, &c->scratch,
                      &c->big_scratch, &c->list, &c->hist, &c->per_root, &c->order, &c->deg, &c->big_list2, &c->spans,
                      &c->batch_off, &c->batch_progs, &c->batch_q, &c->donate, &c->taskbuild, &c->orient, &c->roots})
        release(c, *b);
    if (c->spill) (void)hipHostFree(c->spill);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* g2_last_error(const g2_ctx* c) { return c ? c->err.c_str() : "null context"; }

int g2_upload_view(g2_ctx* c, const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint64_t ne,
                   const uint32_t* src, uint64_t ns) {
    if (!c || !rp || (ne && !cols) || (ns && !src)) return set_err(c, G2_EINVAL, "null argument");
    if (rp[nv] != ne) return set_err(c, G2_EINVAL, "row_ptrs[nv] != ne");
    if (nv >= 0xffffffffull) return set_err(c, G2_EINVAL, "vertex ids must fit in u32");
    for (uint64_t i = 0; i < ns; i++)
        if (src[i] >= nv) return set_err(c, G2_EINVAL, "source id out of range");
    // The kernels index rows and columns without bounds checks, so a malformed
    // CSR is refused here, before anything on the device changes.
    uint64_t md = 0;
    for (uint64_t v = 0; v < nv; v++) {
        if (rp[v + 1] < rp[v]) return set_err(c, G2_EINVAL, "row_ptrs must be non-decreasing");
        md = std::max(md, rp[v + 1] - rp[v]);
    }
    for (uint64_t j = nv ? rp[0] : ne; j < ne; j++)
        if (cols[j] >= nv) return set_err(c, G2_EINVAL, "column id out of range");
    // Level-1 windows of the matching tasks: source i covers windows
    // [off[i], off[i+1]) of gm_chunk row entries each (one window when 0).
    // The local-bitmap matching path needs every neighbour of a source that
    // can root a match to have its row in the view: the largest degree of a
    // source with a rowless neighbour is kept (a view materialised with the
    // plan's hop and root degree leaves such neighbours only next to sources
    // below the root degree, which root no match).
    std::vector<uint64_t> task_off(ns + 1, 0), src_deg(ns);
    const uint32_t chunk = c->tune.gm_chunk;
    uint64_t unclosed = 0;
    for (uint64_t i = 0; i < ns; i++) {
        const uint64_t d = rp[src[i] + 1] - rp[src[i]];
        src_deg[i] = d;
        uint64_t w = 1;
        if (chunk && d > chunk) w = (d + chunk - 1) / chunk;
        task_off[i + 1] = task_off[i] + w;
        if (d > unclosed)
            for (uint64_t j = rp[src[i]]; j < rp[src[i] + 1]; j++)
                if (rp[cols[j] + 1] == rp[cols[j]]) {
                    unclosed = d;
                    break;
                }
    }
    // The budget is checked for every view buffer before any is replaced, so
    // a refused upload leaves the previous view intact.
    struct Need { DevBuf* b; uint64_t bytes; const char* what; };
    const Need needs[] = {{&c->rp, (nv + 1) * 8, "row_ptrs"},
                          {&c->cols, std::max<uint64_t>(ne, 1) * 4, "cols"},
                          {&c->src, std::max<uint64_t>(ns, 1) * 4, "sources"},
                          {&c->deg, (nv + 1) * 4, "degrees"},
                          {&c->gm_task_off, (ns + 1) * 8, "task offsets"}};
    uint64_t after = c->used;
    for (const Need& n : needs) {
        const uint64_t want = (n.bytes + 255) & ~255ull;
        if (want > n.b->bytes) after += want - n.b->bytes;
    }
    if (c->budget && after > c->budget)
        return set_err(c, G2_EBUDGET, "view needs " + std::to_string(after - c->used) + " B more device memory, budget left " +
                                          std::to_string(c->budget - std::min(c->budget, c->used)));
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    // From here on a failure leaves no view: runs return G2_ENOVIEW until an
    // upload succeeds.
    c->view_ok = false;
    c->orient_valid = false;
    c->nv = c->ne = c->ns = 0;
    auto fail = [&](int rc) {
        release(c, c->rp);
        return rc;
    };
    int rc;
    for (const Need& n : needs)
        if ((rc = ensure(c, *n.b, n.bytes, n.what))) return fail(rc);
    std::vector<uint32_t> dg(nv + 1, 0);
    for (uint64_t v = 0; v < nv; v++) dg[v] = (uint32_t)std::min<uint64_t>(rp[v + 1] - rp[v], 0xffffffffull);
    if ((e = hipMemcpyAsync(c->rp.ptr, rp, (nv + 1) * 8, hipMemcpyHostToDevice, c->stream))) return fail(hip_err(c, e, "H2D row_ptrs"));
    if (ne && (e = hipMemcpyAsync(c->cols.ptr, cols, ne * 4, hipMemcpyHostToDevice, c->stream))) return fail(hip_err(c, e, "H2D cols"));
    if (ns && (e = hipMemcpyAsync(c->src.ptr, src, ns * 4, hipMemcpyHostToDevice, c->stream))) return fail(hip_err(c, e, "H2D sources"));
    if ((e = hipMemcpyAsync(c->deg.ptr, dg.data(), (nv + 1) * 4, hipMemcpyHostToDevice, c->stream)))
        return fail(hip_err(c, e, "H2D degrees"));
    if ((e = hipMemcpyAsync(c->gm_task_off.ptr, task_off.data(), (ns + 1) * 8, hipMemcpyHostToDevice, c->stream)))
        return fail(hip_err(c, e, "H2D task offsets"));
    if ((e = hipStreamSynchronize(c->stream))) return fail(hip_err(c, e, "upload sync"));
    c->nv = nv;
    c->ne = ne;
    c->ns = ns;
    c->max_degree = md;
    // whole-graph view: every vertex is a source, in id order
    c->sources_all = (ns == nv);
    for (uint64_t i = 0; c->sources_all && i < ns; i++) c->sources_all = (src[i] == i);
    c->unclosed_max_degree = unclosed;
    c->gm_task_chunk = chunk;
    c->h_gm_task_off.swap(task_off);
    c->h_src_deg.swap(src_deg);
    c->batch_off_valid = false;
    c->view_ok = true;
    return G2_OK;
}

int g2_mce_run(g2_ctx* c, const g2_sched* sched, uint32_t flags, uint64_t* total, uint64_t* hist,
               uint32_t hist_len, uint64_t* per_root_nodes, g2_stats* st) {
    if (!c) return G2_EINVAL;
    if (!c->view_ok) return set_err(c, G2_ENOVIEW, "no view uploaded");
    if (hist && hist_len < 2) return set_err(c, G2_EINVAL, "hist_len must be >= 2");
    if (sched && sched->begin && sched->begin >= sched->end)
        return set_err(c, G2_EINVAL, "empty schedule range");
    if (sched && (sched->begin > sched->end || sched->end > c->ns))
        return set_err(c, G2_EINVAL, "schedule range outside the sources");
    // the permutation itself is checked on the device, while the task list is built
    if (sched && sched->order && sched->n != c->ns) return set_err(c, G2_EINVAL, "schedule length != source count");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    if (st) *st = g2_stats{};
    const int rc = mce_run(c, sched, flags, total, hist, hist_len, per_root_nodes, st);
    if (st && rc == G2_OK) st->device_bytes = c->used;
    return rc;
}

int g2_gm_run_batch(g2_ctx* c, const g2_plan* plans, uint32_t nq, const g2_batch* cfg, uint64_t* counts,
                    uint64_t* nodes, uint32_t* over_budget, g2_stats* st) {
    if (!c) return G2_EINVAL;
    if (!c->view_ok) return set_err(c, G2_ENOVIEW, "no view uploaded");
    if (nq && !plans) return set_err(c, G2_EINVAL, "null plans");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    if (st) *st = g2_stats{};
    const int rc = gm_run_batch(c, plans, nq, cfg, counts, nodes, over_budget, st);
    if (st && rc == G2_OK) st->device_bytes = c->used;
    return rc;
}

int g2_plan_local_eligible(const g2_plan* p) {
    if (!p) return G2_EINVAL;
    g2_ctx tmp;  // host-side plan analysis only: no device state is touched
    if (gm_compile(&tmp, p) != G2_OK) return G2_EINVAL;
    return tmp.local_ok ? 1 : 0;
}

int g2_gm_set_plan(g2_ctx* c, const g2_plan* p) {
    if (!c || !p) return G2_EINVAL;
    return gm_compile(c, p);
}

int g2_gm_run(g2_ctx* c, uint32_t flags, uint64_t* count, uint64_t* per_root_count, g2_stats* st) {
    if (!c) return G2_EINVAL;
    if (!c->view_ok) return set_err(c, G2_ENOVIEW, "no view uploaded");
    if (!c->has_plan) return set_err(c, G2_ENOPLAN, "no plan set");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    if (st) *st = g2_stats{};
    const int rc = gm_run(c, flags, count, per_root_count, st);
    if (st && rc == G2_OK) st->device_bytes = c->used;
    return rc;
}

}  // extern "C"
