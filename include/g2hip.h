/* include/g2hip.h -- device C ABI of the MI355X subgraph-search library
 * (libg2hip.so: gfx950 kernels + device runtime).
 *
 * This is the drop-in boundary that replaces the reference's per-GPU
 * executor and its app hooks:
 *   PipelineExecutor(dev, graph, max_sources, app, reserved, strategy)
 *                                   system/pipeline_executor.h:44-99  -> g2_create
 *   PipelineExecutor::Transfer(holder)  system/pipeline_executor.h:121-149 -> g2_upload_view
 *   PipelineExecutor::Run() with BKExpandSequential
 *                                   system/pipeline_executor.h:151-207,
 *                                   app_BK/BK.h:323-479                 -> g2_mce_run
 *   GMContext / allocateGMMemory    system/appbase.h:11-68,156-216      -> g2_gm_set_plan
 *   PipelineExecutor::Run() with GMatchApp
 *                                   app_gmatch/gmatch.h:67-674          -> g2_gm_run
 *   total_counts_ reduction         system/worker.h:250-255,273-279     (counts are
 *                                   reduced on the device and returned)
 *
 * Only plain integers and pointers cross it.  One context per consumer
 * thread; a context is not thread-safe, several contexts may share a device
 * (each has its own stream and memory budget).
 *
 * Return convention: 0 = OK, < 0 = error (G2_E*); g2_last_error() gives the
 * message.  Host callers restore the reference's behaviour on error (print
 * and exit(-1), device/cuda_context.h:18-25).
 */
#ifndef G2HIP_H
#define G2HIP_H

#include <stdint.h>
#include "g2plan.h"

#ifdef __cplusplus
extern "C" {
#endif

#define G2_OK          0
#define G2_EINVAL     -1   /* bad argument */
#define G2_EHIP       -2   /* HIP runtime error */
#define G2_EBUDGET    -3   /* memory budget too small for the request */
#define G2_ENOVIEW    -4   /* run before upload */
#define G2_ENOPLAN    -5   /* matching run before a plan was set */

typedef struct g2_ctx g2_ctx;

/* launch and scratch parameters (0 = library default) */
typedef struct g2_tuning {
    uint32_t waves_per_cu;        /* resident search waves per CU of the persistent grids
                                     (0 = each kernel's default: 16, local-bitmap matching 32) */
    uint32_t block_waves;         /* waves per workgroup */
    uint64_t wave_scratch_words;  /* per-wave u64 words for MCE local bitsets (small-root pass) */
    uint32_t big_waves;           /* waves of the large-root pass */
    uint32_t gm_chunk;            /* level-1 candidates per matching task (0 = whole root) */
    uint32_t mce_small_max;       /* MCE roots of at most this degree run in registers
                                     (0 = 64, the most; 1 = practically off) */
    uint32_t mce_donate_every;    /* MCE: once a task has walked 64x this many nodes, the wave looks
                                     for idle waves every this many nodes and hands them the upper
                                     half of its shallowest open frame (0 = 256; 0xffffffff = never) */
    uint32_t mce_roots_per_wave;  /* MCE: 2 = roots of degree <= 32 run two per wave (lane groups
                                     of 32), 4 = degree <= 16 four per wave; 0/1 = one per wave */
} g2_tuning;

/* measurements of the last run */
typedef struct g2_stats {
    double   kernel_ms;           /* HIP-event time of the search kernels, all launches (local-bitmap
                                     matching: its counting passes, not the oriented-row build) */
    uint32_t launches;            /* launches of the dominant search kernel */
    double   total_ms;            /* HIP-event time of the whole run on the context's stream */
    uint64_t nodes;               /* MCE search-tree nodes / GM partial-match extensions */
    uint64_t alg_bytes;           /* modelled algorithmic bytes (only with G2_RUN_ACCOUNT) */
    uint64_t big_roots;           /* roots routed to the large-root pass (MCE, canonical matching), or
                                     deferred by the first local-bitmap pass to the workgroup passes */
    uint64_t max_scratch_words;   /* largest per-root scratch need seen */
    double   wave_idle_frac;      /* MCE: 1 - sum of per-wave busy spans / (waves x makespan),
                                     from the GPU real-time clock (tail idleness) */
    uint64_t donations;           /* MCE: search frames handed from busy to idle waves */
    uint32_t local_bitmap;        /* matching: 1 if the oriented local-bitmap passes counted the run */
    uint32_t reserved;
    uint64_t local_pass_roots[5]; /* local-bitmap matching: roots each pass took to the end (rows in
                                     a wave's LDS; in a 256-, 512-, 1024-thread workgroup's LDS;
                                     in per-wave global arenas) */
    uint64_t host_spill_bytes;    /* MCE: page-locked host memory the large-root arenas used because
                                     the device budget could not hold one wave's arena (0 = none) */
    uint64_t device_bytes;        /* device memory the context holds after the run */
    double   local_pass_ms[5];    /* local-bitmap matching: HIP-event time of each pass */
    uint64_t local_pass_bytes[5]; /* with G2_RUN_LOCAL_BYTES: algorithmic bytes of each pass */
} g2_stats;

/* root schedule for g2_mce_run (NULL = the given source order, one wave per root) */
typedef struct g2_sched {
    const uint32_t* order;        /* permutation of [0, ns): processing order of the sources (NULL = as given) */
    uint64_t        n;            /* must equal the uploaded source count when order is set */
    uint64_t        begin, end;   /* process positions [begin, end) of that order only (0,0 = all) */
    const uint32_t* split;        /* per source (source order): deal the root's first-level
                                     children into this many tasks, child c going to task
                                     c mod k (NULL or 0/1 = whole root; at most 65535) */
} g2_sched;

#define G2_RUN_ACCOUNT  1u        /* also compute the modelled algorithmic bytes */
#define G2_RUN_CANONICAL 2u       /* matching: always walk the reference's traversal order
                                     (otherwise whole-graph totals of clique-like plans use
                                     the degree-oriented local-bitmap kernel) */
#define G2_RUN_LOCAL_BYTES 4u     /* matching on the local-bitmap path: also compute the bytes of
                                     oriented rows each counting pass must read (alg_bytes,
                                     local_pass_bytes; an extra untimed kernel) */

int  g2_create(g2_ctx** ctx, int device, uint64_t mem_budget_bytes, const g2_tuning* tuning);
/* Replace the context's launch tuning (0 fields = defaults) between runs, so
 * a tuner can try shapes without re-uploading the view.  gm_chunk keeps the
 * value the last upload used.  No counterpart in the reference. */
int  g2_set_tuning(g2_ctx* ctx, const g2_tuning* tuning);
void g2_destroy(g2_ctx* ctx);
const char* g2_last_error(const g2_ctx* ctx);

/* Copy a view (CSR rows indexed by global vertex id, sorted) and its source
 * vertices to the device.  The caller keeps ownership of the host arrays.
 * G2_EINVAL, with nothing on the device changed, when row_ptrs decreases,
 * row_ptrs[nv] != ne, or a column or source id is >= nv. */
int  g2_upload_view(g2_ctx* ctx, const uint64_t* row_ptrs, uint64_t nv,
                    const uint32_t* cols, uint64_t ne,
                    const uint32_t* sources, uint64_t ns);

/* Maximal clique enumeration rooted at the uploaded sources (pivoted
 * Bron-Kerbosch, app_BK/BK.h).  total = number of maximal cliques;
 * hist[k] = cliques with k vertices (k >= hist_len-1 land in the last bin);
 * per_root_nodes (nullable, ns entries, source order) = search-tree nodes of
 * each root's subtree.  G2_EBUDGET when a root has more than 4096 neighbours
 * above it (P sets are at most 64 words per lane) or its arena does not fit
 * the budget; degeneracy-ordered inputs keep the first bound at the
 * degeneracy. */
int  g2_mce_run(g2_ctx* ctx, const g2_sched* sched, uint32_t flags,
                uint64_t* total, uint64_t* hist, uint32_t hist_len,
                uint64_t* per_root_nodes, g2_stats* stats);

/* Set the query plan for g2_gm_run (tables of system/plan.h). */
int  g2_gm_set_plan(g2_ctx* ctx, const g2_plan* plan);

/* Match count of the plan rooted at the uploaded sources (u64, the
 * overflow-free count of app_gmatch/gmatch.h).  per_root_count nullable.
 * Totals of k-clique-like plans (g2_plan_local_eligible) take the oriented
 * local-bitmap path when the view holds the rows of every neighbour of each
 * source of degree >= n-1 (the whole graph, or a bin materialised with the
 * plan's hop, view/view_bin.h:61-159): a view then counts the matches whose
 * lowest (degree, id) vertex is one of its sources, where the reference's
 * traversal counts those whose lowest-id vertex is; over bins that partition
 * the vertices both give the whole-graph total.  stats->local_bitmap says
 * which path ran: a total summed over views must not mix the two. */
int  g2_gm_run(g2_ctx* ctx, uint32_t flags, uint64_t* count, uint64_t* per_root_count,
               g2_stats* stats);

/* ---- allocation-free matching --------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): g2_gm_run's count of `plan` rooted at `sources`, over a
 * view the caller already holds in device memory (row_ptrs u64[nv+1],
 * cols u32[ne], sources u32[ns], all device pointers, rows sorted), by the
 * canonical traversal (G2_RUN_CANONICAL).  *count is a host pointer.
 * Nothing is allocated: counters and per-wave candidate lists live in a
 * 64 MiB array of the library's own that every device holds until the
 * process ends.  Per wave that takes 8 * n * max(max degree, 64) bytes, so
 * the wave count shrinks with the largest degree; G2_EBUDGET when one wave's
 * lists do not fit (max degree above about 2^23 / n: 2,097,144 for a
 * 4-vertex plan).  G2_EINVAL when the plan is
 * invalid or the CSR is malformed (checked on the device before the search).
 * Runs on the device's null stream and returns when it has finished: the
 * caller's writes to the inputs must be complete before the call.  Calls
 * that overlap on one device share the array and race; calls on different
 * devices do not.  g2_noalloc_last_error() gives the calling thread's last
 * error message. */
int  g2_gm_run_noalloc(int device, const uint64_t* row_ptrs, uint64_t nv, const uint32_t* cols, uint64_t ne,
                       const uint32_t* sources, uint64_t ns, const g2_plan* plan, uint64_t* count);
const char* g2_noalloc_last_error(void);

/* ---- batched matching (no counterpart in the reference: its run binary
 * takes one query per process, app_gmatch/main.cu, system/worker.h:28-280) */

/* launch shape of a batched run */
typedef struct g2_batch {
    const uint32_t* order;        /* processing order of the queries, a permutation of
                                     [0, nq) (NULL = 0..nq-1) */
    uint32_t roots_per_task;      /* root tasks a wave takes per fetch (0 = 1) */
    uint32_t window;              /* level-1 window: candidates of a root row per root task
                                     (0 = whole rows) */
    uint32_t waves_per_cu;        /* resident waves per CU (0 = the context's tuning) */
    uint32_t reserved;
    uint64_t node_budget;         /* per-query limit on partial-match extensions; a query
                                     that passes it is dropped and flagged (0 = none) */
} g2_batch;

/* Count nq plans over the uploaded view in one persistent launch.  counts[q]
 * equals g2_gm_run's count of plans[q] alone (canonical traversal; the same
 * root tasks summed into a per-query u64).  nodes (nullable) receives each
 * query's partial-match extensions; over_budget (nullable) 1 where the query
 * passed cfg->node_budget (its count is then incomplete).  cfg NULL = defaults.
 * stats: kernel_ms of the launch, nodes over all queries. */
int  g2_gm_run_batch(g2_ctx* ctx, const g2_plan* plans, uint32_t nq, const g2_batch* cfg,
                     uint64_t* counts, uint64_t* nodes, uint32_t* over_budget, g2_stats* stats);

/* Autotuner of the batched launch: times g2_gm_run_batch on a seeded sample
 * of at most sample_queries plans (all when 0) by coordinate descent over the
 * level-1 window, roots per task and resident waves per CU, and writes the
 * fastest shape to *best.  best->node_budget is read first and bounds every
 * trial; queries that pass it in the first trial are left out of the timing.
 * best->order is set to NULL.  *best_ms (nullable) = the winner's kernel time
 * on the sample. */
int  g2_gm_autotune(g2_ctx* ctx, const g2_plan* plans, uint32_t nq, uint32_t sample_queries,
                    uint64_t seed, g2_batch* best, double* best_ms);

/* 1 if g2_gm_run counts this plan's whole-graph totals with the oriented
 * local-bitmap kernel: the root is a back neighbour of, and ordered below,
 * every other position, and the order conditions pick exactly one member of
 * every automorphism class for every ranking (so the total does not depend on
 * the vertex order).  0 otherwise; < 0 for an invalid plan.  Host only. */
int  g2_plan_local_eligible(const g2_plan* plan);

/* ---- learned root scheduler (no counterpart in the reference) ----------- */

/* Integer features of every uploaded source, G2_NFEAT_INT per source:
 * degree, the distinct 2-hop ball size |{v} u N(v) u N(N(v))| (HyperLogLog
 * estimate, 64 registers: rounded to the nearest integer, relative standard
 * error about 13%; over a view, N(u) is the row the view holds), triangles at
 * v, neighbours sharing a common neighbour with v, and |{u in N(v): u > v}|
 * (the root's P).  Temporary device memory: 64 B per vertex of the view. */
#define G2_NFEAT_INT 5
int  g2_features(g2_ctx* ctx, int64_t* out);

/* Core number of every vertex of the uploaded view (nv entries), by the
 * h-index iteration on the device; equal to the graph's core numbers for a
 * whole-graph view (a partial view's rows give its own). */
int  g2_core_numbers(g2_ctx* ctx, uint32_t* core);
/* The cost model's 8 inputs of every uploaded source (ns x 8 fp32, source
 * order), computed on the device with one fp32 rounding per operation:
 * log1p(degree), log1p(core number), log1p(2-hop ball), log1p(triangles),
 * local clustering 2t / (d(d-1)), id / nv (degeneracy rank of a recoded
 * graph), shared-neighbour fraction, log1p(later neighbours).  core (host,
 * nv entries) gives the core numbers; NULL computes them from the view
 * (g2_core_numbers).  g2h_root_features is the same formula on the host. */
int  g2_root_inputs(g2_ctx* ctx, const uint32_t* core, float* out);

/* Cost model: MLP 8 -> 64 -> 64 -> 1 (ReLU), BF16 weights/activations on the
 * matrix cores, FP32 accumulation, FP32 master weights, Adam.  Parameters
 * (fp32): W1[8][64], b1[64], W2[64][64], b2[64], W3[64], b3. */
#define G2_MLP_NPARAM 4801
typedef struct g2_mlp g2_mlp;
int  g2_mlp_create(g2_mlp** mlp, int device, uint64_t seed);
void g2_mlp_destroy(g2_mlp* mlp);
int  g2_mlp_get_params(g2_mlp* mlp, float* out);
int  g2_mlp_set_params(g2_mlp* mlp, const float* in);
/* y[i] = f(x[i*8 .. i*8+7]) */
int  g2_mlp_infer(g2_mlp* mlp, const float* x, uint64_t n, float* y);
/* gradient of the mean squared error over n samples (no update) */
int  g2_mlp_grad(g2_mlp* mlp, const float* x, const float* y, uint64_t n, float* grad, float* loss);
/* Adam over seeded shuffled mini-batches; *loss = MSE of the last epoch */
int  g2_mlp_train(g2_mlp* mlp, const float* x, const float* y, uint64_t n, uint32_t epochs,
                  uint32_t batch, float lr, float* loss);

/* The learned schedule of the uploaded view, computed on the device: the
 * features, the 8 inputs standardised as (x - mu[k]) / sd[k] in fp64, the
 * cost model's predictions pred (log(1 + subtree nodes)), then order = the
 * sources longest predicted first (cost = expm1(clamp(pred, 0, 60)), ties
 * in source order) and split = per source ceil(cost / share) parts, at most
 * its later neighbours and max_split (0: 4096), at least 1, with share =
 * max(sum of costs / waves, 1) -- g2h_learned_schedule's rule.  core: host,
 * nv entries (NULL: g2_core_numbers of the view).  order, split (and pred if
 * not NULL): host, ns entries; they feed g2_sched.  mlp must be on the
 * context's device. */
int  g2_learned_schedule(g2_ctx* ctx, g2_mlp* mlp, const float* mu, const float* sd, const uint32_t* core,
                         uint64_t waves, uint32_t max_split, uint32_t* order, uint32_t* split, float* pred);

/* Device properties used for launch sizing */
int  g2_device_info(int device, uint32_t* num_cus, uint64_t* total_mem, char* name, uint32_t name_len);
/* Number of visible GPUs (cudaGetDeviceCount in system/worker.h:45-48). */
int  g2_device_count(void);

/* Page-locked host memory for view holders (cudaMallocHost in
 * view/view_bin_holder.h:28-29); NULL on failure. */
void* g2_pinned_alloc(uint64_t bytes);
void  g2_pinned_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
