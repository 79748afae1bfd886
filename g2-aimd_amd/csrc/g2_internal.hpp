// csrc/g2_internal.hpp -- device-runtime internals shared by the .hip units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "g2hip.h"

namespace g2 {

// device buffer that grows on demand and is charged to the context budget
struct DevBuf {
    void* ptr = nullptr;
    uint64_t bytes = 0;
};

// matching plan as the kernels read it: every table indexed by POSITION in
// the match order, back neighbours and condition operands as positions
struct GmStep {
    uint32_t nbn;                 // back neighbours of this position
    uint32_t bn[G2_MAXQ];
    uint32_t ncond;               // conditions against earlier positions
    uint32_t cop[G2_MAXQ];        // 0: M(u) < M(j)   1: M(u) > M(j)   2: M(u) != M(j)
    uint32_t cpos[G2_MAXQ];
    uint32_t shared;              // candidates come from the pre-intersection
    uint32_t npre, pre[G2_MAXQ];  // pre back neighbours (positions < id-1)
    uint32_t npcond, pcop[G2_MAXQ], pcpos[G2_MAXQ];
    uint32_t nafter, after[G2_MAXQ];
    uint32_t nacond, acop[G2_MAXQ], acpos[G2_MAXQ];
    uint32_t store_prefix;        // strategy[id+1] == PREFIX (byte model only)
    uint32_t pre_is_prev;         // the pre-intersection equals position id-1's candidate set
};

struct GmProgram {
    uint32_t n;
    GmStep step[G2_MAXQ];
};

// canonical matching kernel: counters, then its arguments
enum { GCTR_NEXT = 0, GCTR_COUNT, GCTR_BYTES, GCTR_NODES, GCTR_N };
constexpr int kGmMaxBlockWaves = 4;

struct GmArgs {
    const uint64_t* __restrict__ rp;
    const uint32_t* __restrict__ cols;
    const uint32_t* __restrict__ src;
    const uint64_t* __restrict__ task_off;  // ns+1 window offsets per source (null: one task per source)
    uint64_t ns;
    uint64_t ntasks;
    uint32_t chunk;                         // level-1 window length (0 = whole row)
    uint32_t* scratch;
    uint64_t cap;                           // entries per list
    unsigned long long* ctr;
    unsigned long long* per_root;           // nullable
};

// Oriented local-bitmap form of a plan whose root is adjacent to, and ordered
// below, every other query vertex (all k-cliques among the presets).  Every
// table is by position >= 1; bn_mask excludes the root.
struct LocalProgram {
    uint32_t n;
    uint32_t bn_mask[G2_MAXQ];     // back neighbours (positions >= 1) as a bit mask
    uint32_t gt_mask[G2_MAXQ];     // positions j >= 1 with M(u_i) > M(u_j)
    uint32_t lt_mask[G2_MAXQ];     // positions j >= 1 with M(u_i) < M(u_j)
    uint32_t ne_mask[G2_MAXQ];     // positions j >= 1 with M(u_i) != M(u_j)
    uint32_t clique;               // every position i >= 2 adjacent to 1..i-1 and above i-1, no other masks
};

}  // namespace g2

struct g2_ctx {
    int device = 0;
    uint32_t num_cus = 0;
    uint64_t budget = 0;          // 0 = no limit beyond device memory
    uint64_t used = 0;
    g2_tuning tune{};
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;

    // view
    g2::DevBuf rp, cols, src;
    uint64_t nv = 0, ne = 0, ns = 0;
    uint64_t max_degree = 0;
    bool view_ok = false;         // the last upload completed (runs need a view)
    bool sources_all = false;     // sources == [0, nv) in order (whole-graph view)
    uint64_t unclosed_max_degree = 0;  // largest degree of a source with a rowless neighbour (0 = none)
    g2::DevBuf deg;               // u32 degree of every vertex
    std::vector<uint64_t> h_gm_task_off;  // host copy: level-1 window offsets per source
    g2::DevBuf gm_task_off;
    uint32_t gm_task_chunk = 0;

    // run state
    g2::DevBuf counters;          // small counters / flags
    g2::DevBuf scratch;           // per-wave arenas
    g2::DevBuf big_scratch;
    g2::DevBuf list;              // deferred roots
    g2::DevBuf hist;
    g2::DevBuf per_root;
    g2::DevBuf order;
    g2::DevBuf spans;
    g2::DevBuf donate;            // MCE donation queue (counters, record table, pool)
    g2::DevBuf taskbuild;         // MCE task-list construction scratch
    void* spill = nullptr;        // MCE host spill: page-locked large-root arenas
    uint64_t spill_bytes = 0;

    // matching
    bool has_plan = false;
    g2::GmProgram prog{};
    bool local_ok = false;        // the plan qualifies for the oriented local-bitmap path
    g2::LocalProgram lprog{};
    g2::DevBuf big_list2;
    g2::DevBuf orient;            // oriented relabelled rows of the local-bitmap path
    bool orient_valid = false;    // orient holds the rows of the current view
    g2::DevBuf roots;             // local-bitmap path: ranks of the sources of a partial view
    g2_plan plan{};

    // batched matching
    std::vector<uint64_t> h_src_deg;  // host copy: degree of every source
    g2::DevBuf batch_off;         // task offsets for a batch window != gm_task_chunk
    uint32_t batch_window = 0;
    uint64_t batch_ntasks = 0;
    bool batch_off_valid = false;
    g2::DevBuf batch_progs;       // GmProgram per query
    g2::DevBuf batch_q;           // per-query counts, nodes, order, flags
};

namespace g2 {

int set_err(g2_ctx* c, int code, const std::string& msg);

// resident waves per CU of a persistent grid: the context's setting, else the
// kernel's own default (measured on MI355X: latency-bound searches gain up to
// the occupancy their register use allows)
enum WaveKind { WK_MCE = 16, WK_GM = 16, WK_GM_LOCAL = 32, WK_FEATURES = 16 };
inline uint32_t waves_per_cu(const g2_ctx* c, WaveKind k) {
    return c->tune.waves_per_cu ? c->tune.waves_per_cu : (uint32_t)k;
}
int hip_err(g2_ctx* c, hipError_t e, const char* what);
// make sure buf holds at least `bytes`; charged against the budget
int ensure(g2_ctx* c, DevBuf& buf, uint64_t bytes, const char* what);

int mce_run(g2_ctx* c, const g2_sched* sched, uint32_t flags, uint64_t* total, uint64_t* hist,
            uint32_t hist_len, uint64_t* per_root_nodes, g2_stats* st);
int gm_run(g2_ctx* c, uint32_t flags, uint64_t* count, uint64_t* per_root_count, g2_stats* st);
int gm_compile(g2_ctx* c, const g2_plan* p);
// plan tables re-indexed by position; G2_EINVAL with *why naming the first broken table
int gm_program(const g2_plan* p, GmProgram* out, const char** why);
// gm_kernel (no byte accounting) over `blocks` workgroups of block_waves waves
hipError_t gm_launch(const GmArgs& a, const GmProgram& prog, uint64_t blocks, uint32_t block_waves, hipStream_t s);
void gm_local_compile(g2_ctx* c, const g2_plan* p);
// triangle count per rank (tri, u64) and triangle-carrying oriented entries
// per rank (sup, u32) of the uploaded view, through its oriented rows (built
// and cached as for local matching); *rank = the rank of every vertex
int local_triangle_features(g2_ctx* c, unsigned long long* tri, uint32_t* sup, uint8_t* flag, uint64_t flag_bytes,
                            const uint32_t** rank);
// drop the cached oriented rows (their memory returns to the budget)
void orient_release(g2_ctx* c);
int gm_local_run(g2_ctx* c, uint64_t* count, g2_stats* st, bool bytes);
int gm_run_batch(g2_ctx* c, const g2_plan* plans, uint32_t nq, const g2_batch* cfg, uint64_t* counts,
                 uint64_t* qnodes, uint32_t* over, g2_stats* st);

}  // namespace g2
