// csrc/g2_gm.hip -- plan-driven subgraph matching on gfx950.
//
// Counts the same matches as the reference's GMatch app
// (app_gmatch/gmatch.h:83-674): a partial match of length id is extended by
// position id of the plan's match order with
//   candidates = row of the back neighbour with the smallest row (first on
//                ties, gmatch.h:517-536), filtered by the conditions against
//                earlier positions (gmatch.h:559-588) and by membership in the
//                rows of the other back neighbours (gmatch.h:589-601);
//   shared positions take a pre-intersection over the pre back neighbours,
//                computed once per prefix and filtered per candidate of the
//                previous position by the after conditions / rows
//                (gmatch.h:273-509);
//   the last position is counted, not stored (gmatch.h:604-608).
// Counts are the overflow-free u64 totals.
//
// MI355X design: one wave64 walks one root (or one window of the root's row
// at position 1) depth first.  Each level keeps its candidate list in the
// wave's HBM arena (written by ballot compaction, read back as broadcasts);
// order conditions become a [lo,hi) value range so only the qualifying slice
// of every sorted row is streamed or searched.  The last level counts with a
// per-lane u64 and the root total is one wave reduction.
#include <algorithm>

#include "g2_device.hpp"
#include "g2_internal.hpp"

namespace g2 {



struct Range {
    uint32_t lo;   // smallest admissible value
    uint32_t hi;   // one past the largest admissible value (0xffffffff = none)
};


struct WaveState {
    uint32_t M[G2_MAXQ];
    uint64_t len[G2_MAXQ];
    uint64_t idx[G2_MAXQ];
    uint64_t prelen[G2_MAXQ];
    uint64_t olo[G2_MAXQ], ohi[G2_MAXQ];  // narrowed rows of the membership tests (wave-uniform)
};

__device__ inline Range cond_range(const WaveState& S, uint32_t n, const uint32_t* op, const uint32_t* pos) {
    Range r{0u, 0xffffffffu};
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t m = S.M[pos[k]];
        if (op[k] == 1) r.lo = max(r.lo, m + 1);          // M(u) > M(j)
        else if (op[k] == 0) r.hi = min(r.hi, m);          // M(u) < M(j)
    }
    return r;
}

__device__ inline bool neq_ok(const WaveState& S, uint32_t n, const uint32_t* op, const uint32_t* pos, uint32_t w) {
    for (uint32_t k = 0; k < n; k++)
        if (op[k] == 2 && S.M[pos[k]] == w) return false;
    return true;
}

__device__ inline uint64_t store_bytes(const GmStep& st, uint32_t id, uint32_t n, uint64_t nout) {
    if (id + 1 == n) return 0;
    if (st.store_prefix) return 4 * (id + nout);
    return 4 * (uint64_t)(id + 1) * nout;
}

// Filter the row of `parent` (restricted to [row_lo,row_hi) entries) by the
// value range, NON_EQUAL conditions and membership in the other rows.  With
// out != nullptr the survivors are appended (sorted) and their number is
// returned; otherwise they are only counted into *lane_count (per lane).
template <bool ACCOUNT>
__device__ uint64_t filter_row(const GmArgs& a, WaveState& S, uint32_t nbn, const uint32_t* bn,
                               uint32_t parent, uint64_t row_lo, uint64_t row_hi, Range rg,
                               uint32_t ncond, const uint32_t* op, const uint32_t* pos, uint32_t* out,
                               uint64_t* lane_count, uint64_t* bytes, uint64_t* npass_out) {
    const uint32_t lane = lane_id();
    const uint32_t* __restrict__ cols = a.cols;
    uint64_t s0 = lower_bound(cols, row_lo, row_hi, rg.lo);
    uint64_t s1 = rg.hi == 0xffffffffu ? row_hi : lower_bound(cols, s0, row_hi, rg.hi);
    // narrowed rows of the other back neighbours
    // narrowed rows of the other back neighbours, kept in LDS: the values are
    // wave-uniform and an indexed register array would spill
    uint32_t no = 0;
    uint64_t bcost = 0;
    for (uint32_t k = 0; k < nbn; k++) {
        if (bn[k] == parent) continue;
        const uint32_t m = S.M[bn[k]];
        uint64_t l = a.rp[m], h = a.rp[m + 1];
        if (ACCOUNT) bcost += cost_isect(1, h - l);
        l = lower_bound(cols, l, h, rg.lo);
        if (rg.hi != 0xffffffffu) h = lower_bound(cols, l, h, rg.hi);
        if (lane == 0) {
            S.olo[no] = l;
            S.ohi[no] = h;
        }
        no++;
    }
    wave_sync();
    if (!out && no == 0) {
        // counting with no membership test: the whole slice passes except the
        // entries a NON_EQUAL condition names (matched values are distinct)
        uint64_t cnt = s1 - s0;
        for (uint32_t k = 0; k < ncond; k++) {
            const uint32_t m = S.M[pos[k]];
            if (op[k] == 2 && m >= rg.lo && m < rg.hi && contains(cols, s0, s1, m)) cnt--;
        }
        if (lane == 0) *lane_count += cnt;
        if (ACCOUNT && npass_out) *npass_out = cnt;
        return 0;
    }
    uint64_t nout = 0, npass = 0;
    for (uint64_t base = s0; base < s1; base += kWave) {
        const uint64_t e = base + lane;
        bool ok = false, pass = false;
        uint32_t w = 0;
        if (e < s1) {
            w = cols[e];
            pass = neq_ok(S, ncond, op, pos, w);
            ok = pass;
            for (uint32_t k = 0; k < no && ok; k++) ok = contains(cols, S.olo[k], S.ohi[k], w);
        }
        if (ACCOUNT) npass += pass;
        const uint64_t m = ballot(ok);
        if (out) {
            if (ok) out[nout + lanes_below(m)] = w;
            nout += __popcll(m);
        } else if (ok) {
            (*lane_count)++;
        }
    }
    if (ACCOUNT) {
        npass = wave_sum(npass);
        if (lane == 0) *bytes += npass * bcost;
        if (npass_out) *npass_out = npass;
    }
    return nout;
}

// count or collect the after-filtered entries of a pre list for position id
template <bool ACCOUNT>
__device__ uint64_t filter_pre(const GmArgs& a, WaveState& S, const GmStep& st, const uint32_t* pre,
                               uint64_t plen, uint32_t* out, uint64_t* lane_count, uint64_t* bytes) {
    const uint32_t lane = lane_id();
    const Range rg = cond_range(S, st.nacond, st.acop, st.acpos);
    // pre is sorted: restrict to the value range
    const uint64_t p0 = lower_bound(pre, 0, plen, rg.lo);
    const uint64_t p1 = rg.hi == 0xffffffffu ? plen : lower_bound(pre, p0, plen, rg.hi);
    for (uint32_t k = 0; k < st.nafter; k++) {
        const uint32_t m = S.M[st.after[k]];
        uint64_t l = a.rp[m], h = a.rp[m + 1];
        if (ACCOUNT && lane == 0) *bytes += cost_isect(plen, h - l);
        l = lower_bound(a.cols, l, h, rg.lo);
        if (rg.hi != 0xffffffffu) h = lower_bound(a.cols, l, h, rg.hi);
        if (lane == 0) {
            S.olo[k] = l;
            S.ohi[k] = h;
        }
    }
    wave_sync();
    if (!out && st.nafter == 0) {  // counting, no after rows: closed form as in filter_row
        uint64_t cnt = p1 - p0;
        for (uint32_t k = 0; k < st.nacond; k++) {
            const uint32_t m = S.M[st.acpos[k]];
            if (st.acop[k] == 2 && m >= rg.lo && m < rg.hi && contains(pre, p0, p1, m)) cnt--;
        }
        if (lane == 0) *lane_count += cnt;
        return 0;
    }
    uint64_t nout = 0;
    for (uint64_t base = p0; base < p1; base += kWave) {
        const uint64_t e = base + lane;
        bool ok = false;
        uint32_t w = 0;
        if (e < p1) {
            w = pre[e];
            ok = neq_ok(S, st.nacond, st.acop, st.acpos, w);
            for (uint32_t k = 0; k < st.nafter && ok; k++) ok = contains(a.cols, S.olo[k], S.ohi[k], w);
        }
        const uint64_t m = ballot(ok);
        if (out) {
            if (ok) out[nout + lanes_below(m)] = w;
            nout += __popcll(m);
        } else if (ok) {
            (*lane_count)++;
        }
    }
    return nout;
}

// parent = back neighbour whose matched row is smallest, first on ties
__device__ inline uint32_t min_row(const GmArgs& a, const WaveState& S, uint32_t nbn, const uint32_t* bn) {
    uint32_t best = bn[0];
    uint64_t bd = a.rp[S.M[bn[0]] + 1] - a.rp[S.M[bn[0]]];
    for (uint32_t k = 1; k < nbn; k++) {
        const uint32_t m = S.M[bn[k]];
        const uint64_t d = a.rp[m + 1] - a.rp[m];
        if (d < bd) { bd = d; best = bn[k]; }
    }
    return best;
}

// Per-query node budget of the batched launch: a wave charges its nodes to
// the query's counter in blocks of kBudgetBlock; once the total passes the
// limit the query's flag is raised and every wave working on it drops it.
struct NodeBudget {
    unsigned long long* qnodes;
    unsigned int* qflag;
    uint64_t limit;
};

constexpr uint64_t kBudgetBlock = 256;

// charge `pending` nodes of query q; true when the query is over budget
__device__ inline bool budget_charge(const NodeBudget& nb, uint32_t q, uint64_t& pending) {
    __builtin_amdgcn_wave_barrier();  // converged here (see fetch_task)
    uint32_t over = 0;
    if (lane_id() == 0) {
        const unsigned long long old = atomicAdd(&nb.qnodes[q], (unsigned long long)pending);
        if (old + pending > nb.limit) atomicOr(&nb.qflag[q], 1u);
        over = __hip_atomic_load(&nb.qflag[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    pending = 0;
    return uniform(over) != 0;
}

// One root task: source r, level-1 window `win`, under program P.  The
// wave's lists live in `ws` (L(id) at id*cap, PRE(id) at (levels+id)*cap).
// Returns false when the query's node budget ran out (BUDGET only).
template <bool ACCOUNT, bool BUDGET>
__device__ bool gm_root(const GmArgs& a, const GmProgram& P, WaveState& S, uint32_t* ws, uint32_t levels,
                        uint64_t r, uint64_t win, uint64_t& lane_total, uint64_t& bytes, uint64_t& nodes,
                        const NodeBudget& nb, uint32_t q, uint64_t& pending) {
    const uint32_t lane = lane_id();
    const uint32_t n = P.n;
    const uint64_t cap = a.cap;
    auto L = [&](uint32_t id) { return ws + (uint64_t)id * cap; };
    auto PRE = [&](uint32_t id) { return ws + (uint64_t)(levels + id) * cap; };
    const uint32_t v0 = uniform(a.src[r]);
    const uint64_t rlo = a.rp[v0], rhi = a.rp[v0 + 1];
    uint64_t wlo = rlo, whi = rhi;
    if (a.chunk) {
        wlo = min(rhi, rlo + win * a.chunk);
        whi = min(rhi, wlo + a.chunk);
    }
    if (lane == 0) S.M[0] = v0;
    wave_sync();
    if (ACCOUNT && win == 0 && lane == 0) bytes += 4 + 4 * 1 + 4 * (rhi - rlo);

    // position 1: the root's row (window)
    {
        const GmStep& st = P.step[1];
        const Range rg = cond_range(S, st.ncond, st.cop, st.cpos);
        uint64_t nout = 0, bsum = 0, npass = 0;
        nodes++;
        if (BUDGET) pending++;
        if (n == 2) {
            filter_row<ACCOUNT>(a, S, st.nbn, st.bn, 0, wlo, whi, rg, st.ncond, st.cop, st.cpos, nullptr,
                                &lane_total, &bsum, &npass);
        } else {
            nout = filter_row<ACCOUNT>(a, S, st.nbn, st.bn, 0, wlo, whi, rg, st.ncond, st.cop, st.cpos, L(1),
                                       nullptr, &bsum, &npass);
            if (ACCOUNT && lane == 0) {
                bsum += store_bytes(st, 1, n, nout);
                // a PREFIX store writes the prefix once per root, not per window
                if (win && st.store_prefix) bsum -= 4;
            }
            if (lane == 0) { S.len[1] = nout; S.idx[1] = 0; }
        }
        if (ACCOUNT && lane == 0) bytes += bsum;
    }
    wave_sync();
    if (n <= 2) return true;
    // pre-intersection of position 2 if shared (depends on position 0 only)
    uint32_t depth = 1;
    auto build_pre = [&](uint32_t id, bool charge) {
        // prefix = positions < id-1 are matched
        const GmStep& st = P.step[id];
        if (!ACCOUNT && st.pre_is_prev) {  // PRE(id) is L(id-1), already built
            wave_sync();
            return;
        }
        const uint32_t parent = min_row(a, S, st.npre, st.pre);
        const uint32_t pm = S.M[parent];
        const Range rg = cond_range(S, st.npcond, st.pcop, st.pcpos);
        uint64_t bsum = 0;
        const uint64_t len = filter_row<ACCOUNT>(a, S, st.npre, st.pre, parent, a.rp[pm], a.rp[pm + 1], rg,
                                                 st.npcond, st.pcop, st.pcpos, PRE(id), nullptr, &bsum, nullptr);
        if (lane == 0) {
            S.prelen[id] = len;
            if (ACCOUNT && charge) bytes += bsum + 4 * (id - 1) + 4 * (a.rp[pm + 1] - a.rp[pm]);
        }
        wave_sync();
    };
    if (P.step[2].shared) build_pre(2, win == 0);  // same list for every window
    while (true) {
        const uint32_t id = depth;
        const uint64_t ix = S.idx[id];
        if (ix >= S.len[id]) {
            if (depth == 1) break;
            depth--;
            continue;
        }
        const uint32_t cand = L(id)[ix];
        if (lane == 0) { S.M[id] = cand; S.idx[id] = ix + 1; }
        wave_sync();
        const uint32_t nx = id + 1;
        const GmStep& st = P.step[nx];
        nodes++;
        if (BUDGET && ++pending >= kBudgetBlock && budget_charge(nb, q, pending)) return false;
        uint64_t bsum = 0;
        if (st.shared) {
            const bool alias = !ACCOUNT && st.pre_is_prev;
            const uint32_t* pl = alias ? L(id) : PRE(nx);
            const uint64_t plen = alias ? S.len[id] : S.prelen[nx];
            if (nx + 1 == n) {
                filter_pre<ACCOUNT>(a, S, st, pl, plen, nullptr, &lane_total, &bsum);
            } else {
                const uint64_t nout = filter_pre<ACCOUNT>(a, S, st, pl, plen, L(nx), nullptr, &bsum);
                if (ACCOUNT && lane == 0) bsum += store_bytes(st, nx, n, nout);
                if (lane == 0) { S.len[nx] = nout; S.idx[nx] = 0; }
            }
        } else {
            const uint32_t parent = min_row(a, S, st.nbn, st.bn);
            const uint32_t pm = S.M[parent];
            const Range rg = cond_range(S, st.ncond, st.cop, st.cpos);
            if (ACCOUNT && lane == 0) bsum += 4 * nx + 4 * (a.rp[pm + 1] - a.rp[pm]);
            if (nx + 1 == n) {
                filter_row<ACCOUNT>(a, S, st.nbn, st.bn, parent, a.rp[pm], a.rp[pm + 1], rg, st.ncond, st.cop,
                                    st.cpos, nullptr, &lane_total, &bsum, nullptr);
            } else {
                const uint64_t nout = filter_row<ACCOUNT>(a, S, st.nbn, st.bn, parent, a.rp[pm], a.rp[pm + 1],
                                                          rg, st.ncond, st.cop, st.cpos, L(nx), nullptr, &bsum,
                                                          nullptr);
                if (ACCOUNT && lane == 0) bsum += store_bytes(st, nx, n, nout);
                if (lane == 0) { S.len[nx] = nout; S.idx[nx] = 0; }
            }
        }
        if (ACCOUNT && lane == 0) bytes += bsum;
        wave_sync();
        if (nx + 1 < n) {
            if (P.step[nx + 1].shared) build_pre(nx + 1, true);
            depth = nx;
        }
    }
    return true;
}

// task t of a query -> (source, level-1 window)
__device__ inline uint64_t task_source(const GmArgs& a, uint64_t t, uint64_t* win) {
    if (!a.task_off) {  // whole rows: task t is source t
        *win = 0;
        return t;
    }
    uint64_t lo = 0, hi = a.ns;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a.task_off[mid] <= t) lo = mid; else hi = mid;
    }
    const uint64_t r = uniform(lo);
    *win = t - a.task_off[r];
    return r;
}

template <bool ACCOUNT>
__global__ __launch_bounds__(256) void gm_kernel(GmArgs a, GmProgram prog) {
    __shared__ WaveState states[kGmMaxBlockWaves];
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x / kWave;
    WaveState& S = states[wib];
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + wib;
    uint32_t* ws = a.scratch + gwave * a.cap * (2 * prog.n);
    uint64_t lane_total = 0, bytes = 0, nodes = 0, pending = 0;
    const NodeBudget nb{nullptr, nullptr, 0};

    while (true) {
        const uint64_t t = fetch_task(&a.ctr[GCTR_NEXT]);
        if (t >= a.ntasks) break;
        uint64_t win;
        const uint64_t r = task_source(a, t, &win);
        const uint64_t before = lane_total;
        gm_root<ACCOUNT, false>(a, prog, S, ws, prog.n, r, win, lane_total, bytes, nodes, nb, 0, pending);
        if (a.per_root) {
            const uint64_t mine = wave_sum(lane_total - before);
            if (lane == 0 && mine) atomicAdd(&a.per_root[r], (unsigned long long)mine);
        }
    }
    const uint64_t tot = wave_sum(lane_total);
    if (ACCOUNT) bytes = wave_sum(bytes);
    if (lane == 0) {
        if (tot) atomicAdd(&a.ctr[GCTR_COUNT], (unsigned long long)tot);
        if (ACCOUNT && bytes) atomicAdd(&a.ctr[GCTR_BYTES], (unsigned long long)bytes);
        atomicAdd(&a.ctr[GCTR_NODES], (unsigned long long)nodes);
    }
}

// Many plans in one persistent launch.  Task t is query order[t / per_q]'s
// block t % per_q of `rpt` consecutive root tasks; its matches go to that
// query's counter, so each count is the sum over the same root tasks the
// solo launch runs.
struct GmBatchArgs {
    const GmProgram* __restrict__ progs;
    const uint32_t* __restrict__ order;
    uint32_t nq;
    uint32_t levels;             // list levels per wave arena (largest plan size)
    uint32_t rpt;                // root tasks per batch task
    uint64_t per_q;              // batch tasks per query
    unsigned long long* qcount;
    NodeBudget nb;               // qnodes always counted; limit ~0 = none
};

__global__ __launch_bounds__(256) void gm_batch_kernel(GmArgs a, GmBatchArgs b) {
    __shared__ WaveState states[kGmMaxBlockWaves];
    const uint32_t lane = lane_id();
    const uint32_t wib = threadIdx.x / kWave;
    WaveState& S = states[wib];
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + wib;
    uint32_t* ws = a.scratch + gwave * a.cap * (2 * b.levels);
    const uint64_t total = (uint64_t)b.nq * b.per_q;
    uint64_t nodes_all = 0, bytes = 0;

    while (true) {
        const uint64_t t = fetch_task(&a.ctr[GCTR_NEXT]);
        if (t >= total) break;
        const uint32_t q = uniform(b.order[t / b.per_q]);
        uint32_t over = 0;
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) over = __hip_atomic_load(&b.nb.qflag[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (uniform(over)) continue;
        const GmProgram& P = b.progs[q];
        const uint64_t t0 = (t % b.per_q) * b.rpt;
        const uint64_t t1 = min(a.ntasks, t0 + b.rpt);
        uint64_t lane_total = 0, nodes = 0, pending = 0;
        bool ok = true;
        for (uint64_t rt = t0; rt < t1 && ok; rt++) {
            uint64_t win;
            const uint64_t r = task_source(a, rt, &win);
            ok = gm_root<false, true>(a, P, S, ws, b.levels, r, win, lane_total, bytes, nodes, b.nb, q, pending);
        }
        if (ok && pending) budget_charge(b.nb, q, pending);
        nodes_all += nodes;
        const uint64_t tot = wave_sum(lane_total);
        if (lane == 0 && tot) atomicAdd(&b.qcount[q], (unsigned long long)tot);
    }
    if (lane == 0) atomicAdd(&a.ctr[GCTR_NODES], (unsigned long long)nodes_all);
}

// Largest grid (at most nblocks) whose wave arenas fit the context budget,
// or, without one, half of the device memory not yet in use (hub rows make
// max-degree arenas large: R-MAT scale 24 has rows of ~10^5-10^6 entries).
static uint64_t fit_arenas(g2_ctx* c, uint64_t nblocks, uint64_t per_block) {
    uint64_t cap = c->budget ? c->budget - std::min(c->budget, c->used - c->scratch.bytes) : 0;
    if (!c->budget) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) cap = (free_b + c->scratch.bytes) / 2;
    }
    while (nblocks > 1 && cap && nblocks * per_block > cap) nblocks /= 2;
    return nblocks;
}

// Plan tables (system/plan.h) re-indexed by position.  Returns G2_OK, or
// G2_EINVAL with *why naming the first broken table.
int gm_program(const g2_plan* p, GmProgram* out, const char** why) {
    if (p->n < 2 || p->n > G2_MAXQ) return (*why = "plan size must be 2..8", G2_EINVAL);
    const uint32_t n = p->n;
    // Every table below is indexed through seq/rev and sized G2_MAXQ: a plan
    // that breaks these bounds is refused before any entry is read through it.
    for (uint32_t i = 0; i < n; i++)
        if (p->seq[i] >= n || p->rev[p->seq[i]] != i)
            return (*why = "seq and rev are not inverse permutations", G2_EINVAL);
    auto vertices_ok = [n](uint32_t cnt, const uint32_t* v) {
        if (cnt > G2_MAXQ) return false;
        for (uint32_t k = 0; k < cnt; k++)
            if (v[k] >= n) return false;
        return true;
    };
    auto conds_ok = [n](uint32_t cnt, const uint32_t (*cd)[2]) {
        if (cnt > G2_MAXQ) return false;
        for (uint32_t k = 0; k < cnt; k++)
            if (cd[k][1] >= n) return false;
        return true;
    };
    GmProgram g{};
    g.n = n;
    for (uint32_t id = 1; id < n; id++) {
        GmStep& s = g.step[id];
        const uint32_t u = p->seq[id];
        if (p->bn_count[id] == 0 || p->bn_count[id] > G2_MAXQ) return (*why = "position without back neighbour", G2_EINVAL);
        if (!conds_ok(p->cond_count[u], p->cond[u])) return (*why = "bad condition", G2_EINVAL);
        if (p->shared[id] &&
            (!vertices_ok(p->pre_bn_count[id], p->pre_bn[id]) || !vertices_ok(p->after_bn_count[id], p->after_bn[id]) ||
             !conds_ok(p->pre_cond_count[u], p->pre_cond[u]) || !conds_ok(p->after_cond_count[u], p->after_cond[u])))
            return (*why = "shared position tables out of range", G2_EINVAL);
        s.nbn = p->bn_count[id];
        for (uint32_t k = 0; k < s.nbn; k++) {
            const uint32_t b = p->bn[id][k];
            if (b >= n || p->rev[b] >= id) return (*why = "back neighbour not matched earlier", G2_EINVAL);
            s.bn[k] = p->rev[b];
        }
        s.ncond = p->cond_count[u];
        for (uint32_t k = 0; k < s.ncond; k++) {
            s.cop[k] = p->cond[u][k][0];
            s.cpos[k] = p->rev[p->cond[u][k][1]];
            if (s.cop[k] > 2 || s.cpos[k] >= id) return (*why = "bad condition", G2_EINVAL);
        }
        s.shared = p->shared[id] ? 1 : 0;
        if (s.shared) {
            if (id < 2 || p->pre_bn_count[id] == 0) return (*why = "shared position without pre back neighbours", G2_EINVAL);
            s.npre = p->pre_bn_count[id];
            for (uint32_t k = 0; k < s.npre; k++) {
                s.pre[k] = p->rev[p->pre_bn[id][k]];
                if (s.pre[k] + 1 >= id) return (*why = "pre back neighbour not in the prefix", G2_EINVAL);
            }
            s.nafter = p->after_bn_count[id];
            for (uint32_t k = 0; k < s.nafter; k++) {
                s.after[k] = p->rev[p->after_bn[id][k]];
                if (s.after[k] + 1 != id) return (*why = "after back neighbour is not the previous position", G2_EINVAL);
            }
            s.npcond = p->pre_cond_count[u];
            for (uint32_t k = 0; k < s.npcond; k++) {
                s.pcop[k] = p->pre_cond[u][k][0];
                s.pcpos[k] = p->rev[p->pre_cond[u][k][1]];
                if (s.pcop[k] > 2 || s.pcpos[k] + 1 >= id) return (*why = "bad pre condition", G2_EINVAL);
            }
            s.nacond = p->after_cond_count[u];
            for (uint32_t k = 0; k < s.nacond; k++) {
                s.acop[k] = p->after_cond[u][k][0];
                s.acpos[k] = p->rev[p->after_cond[u][k][1]];
                if (s.acop[k] > 2 || s.acpos[k] + 1 != id) return (*why = "bad after condition", G2_EINVAL);
            }
        }
        s.store_prefix = (id + 1 < n && p->strategy[id + 1] == 1) ? 1 : 0;
    }
    if (g.step[1].shared) return (*why = "position 1 cannot be shared", G2_EINVAL);
    // A shared position whose pre back neighbours and pre conditions are, as
    // sets, the back neighbours and conditions of the previous (non-shared)
    // position filters exactly the same candidates: reuse that list.
    for (uint32_t id = 2; id < n; id++) {
        GmStep& s = g.step[id];
        const GmStep& prev = g.step[id - 1];
        if (!s.shared || prev.shared || s.npre != prev.nbn || s.npcond != prev.ncond) continue;
        uint32_t mb = 0, mp = 0;
        for (uint32_t k = 0; k < s.npre; k++) mb |= 1u << s.pre[k];
        for (uint32_t k = 0; k < prev.nbn; k++) mp |= 1u << prev.bn[k];
        if (mb != mp) continue;
        bool same = true;
        for (uint32_t k = 0; k < s.npcond && same; k++) {
            bool found = false;
            for (uint32_t l = 0; l < prev.ncond; l++)
                found |= prev.cop[l] == s.pcop[k] && prev.cpos[l] == s.pcpos[k];
            same = found;
        }
        s.pre_is_prev = same ? 1 : 0;
    }
    *out = g;
    return G2_OK;
}

int gm_compile(g2_ctx* c, const g2_plan* p) {
    GmProgram g{};
    const char* why = "";
    if (gm_program(p, &g, &why) != G2_OK) return set_err(c, G2_EINVAL, why);
    c->prog = g;
    c->plan = *p;
    c->has_plan = true;
    gm_local_compile(c, p);
    return G2_OK;
}

int gm_run(g2_ctx* c, uint32_t flags, uint64_t* count, uint64_t* per_root, g2_stats* st) {
    const bool account = flags & G2_RUN_ACCOUNT;
    // totals of a root-dominated plan over the whole graph or a view in
    // which every source that can root a match (degree >= n-1) has the rows
    // of all its neighbours: oriented local bitmaps (per-root counts and the
    // byte model follow the reference's traversal)
    // (it keeps row offsets in 32 bits and sorts ranks as int-sized lists: fewer
    // than 2^32 row entries and 2^31 - 1 vertices)
    const bool closed = c->sources_all || c->unclosed_max_degree + 1 < (uint64_t)c->lprog.n;
    if (c->local_ok && closed && c->ne < (1ull << 32) && c->nv + 1 <= 0x7fffffffull && !per_root &&
        !account && !(flags & G2_RUN_CANONICAL))
        return gm_local_run(c, count, st, flags & G2_RUN_LOCAL_BYTES);
    const uint32_t bw = std::min<uint32_t>(c->tune.block_waves, kGmMaxBlockWaves);
    uint64_t nblocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * waves_per_cu(c, WK_GM) / bw);
    const uint64_t cap = std::max<uint64_t>(c->max_degree, 64);
    const uint64_t per_wave = cap * 2 * c->prog.n * 4;
    // shrink the grid until the arenas fit the budget
    nblocks = fit_arenas(c, nblocks, bw * per_wave);
    int rc;
    if ((rc = ensure(c, c->counters, GCTR_N * 8, "counters"))) return rc;
    if ((rc = ensure(c, c->scratch, nblocks * bw * per_wave, "wave arenas"))) return rc;
    if (per_root && (rc = ensure(c, c->per_root, std::max<uint64_t>(c->ns, 1) * 8, "per-root counts"))) return rc;
    hipError_t e;
    auto* ctr = (unsigned long long*)c->counters.ptr;
    if ((e = hipMemsetAsync(ctr, 0, GCTR_N * 8, c->stream))) return hip_err(c, e, "memset counters");
    if (per_root && (e = hipMemsetAsync(c->per_root.ptr, 0, c->ns * 8, c->stream))) return hip_err(c, e, "memset per-root");
    GmArgs a{};
    a.rp = (const uint64_t*)c->rp.ptr;
    a.cols = (const uint32_t*)c->cols.ptr;
    a.src = (const uint32_t*)c->src.ptr;
    a.task_off = (const uint64_t*)c->gm_task_off.ptr;
    a.ns = c->ns;
    a.ntasks = c->h_gm_task_off.empty() ? 0 : c->h_gm_task_off.back();
    a.chunk = c->gm_task_chunk;
    a.scratch = (uint32_t*)c->scratch.ptr;
    a.cap = cap;
    a.ctr = ctr;
    a.per_root = per_root ? (unsigned long long*)c->per_root.ptr : nullptr;
    if ((e = hipEventRecord(c->ev[0], c->stream))) return hip_err(c, e, "event");
    if (account) hipLaunchKernelGGL(gm_kernel<true>, dim3(nblocks), dim3(bw * kWave), 0, c->stream, a, c->prog);
    else hipLaunchKernelGGL(gm_kernel<false>, dim3(nblocks), dim3(bw * kWave), 0, c->stream, a, c->prog);
    if ((e = hipGetLastError())) return hip_err(c, e, "gm launch");
    if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
    unsigned long long h[GCTR_N];
    if ((e = hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H counters");
    if (per_root && c->ns && (e = hipMemcpyAsync(per_root, c->per_root.ptr, c->ns * 8, hipMemcpyDeviceToHost, c->stream)))
        return hip_err(c, e, "D2H per-root");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "gm sync");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    if (count) *count = h[GCTR_COUNT];
    if (st) {
        st->kernel_ms = ms;
        st->launches = 1;
        st->total_ms = ms;
        st->nodes = h[GCTR_NODES];
        st->alg_bytes = account ? h[GCTR_BYTES] : 0;
        st->big_roots = 0;
        st->max_scratch_words = cap;
    }
    return G2_OK;
}

}  // namespace g2

namespace g2 {

// level-1 window offsets of the sources for window length w (0 = one task per source)
static int batch_task_offsets(g2_ctx* c, uint32_t w, const uint64_t** dev_off, uint64_t* ntasks) {
    if (w == c->gm_task_chunk) {
        *dev_off = (const uint64_t*)c->gm_task_off.ptr;
        *ntasks = c->h_gm_task_off.back();
        return G2_OK;
    }
    if (!c->batch_off_valid || c->batch_window != w) {
        std::vector<uint64_t> off(c->ns + 1, 0);
        for (uint64_t i = 0; i < c->ns; i++) {
            const uint64_t d = c->h_src_deg[i];
            off[i + 1] = off[i] + ((w && d > w) ? (d + w - 1) / w : 1);
        }
        int rc;
        if ((rc = ensure(c, c->batch_off, (c->ns + 1) * 8, "batch task offsets"))) return rc;
        hipError_t e = hipMemcpy(c->batch_off.ptr, off.data(), (c->ns + 1) * 8, hipMemcpyHostToDevice);
        if (e) return hip_err(c, e, "H2D batch task offsets");
        c->batch_window = w;
        c->batch_ntasks = off.back();
        c->batch_off_valid = true;
    }
    *dev_off = (const uint64_t*)c->batch_off.ptr;
    *ntasks = c->batch_ntasks;
    return G2_OK;
}

int gm_run_batch(g2_ctx* c, const g2_plan* plans, uint32_t nq, const g2_batch* cfg, uint64_t* counts,
                 uint64_t* qnodes, uint32_t* over, g2_stats* st) {
    const g2_batch dflt{};
    if (!cfg) cfg = &dflt;
    if (st) *st = g2_stats{};
    if (nq == 0) return G2_OK;
    std::vector<GmProgram> progs(nq);
    uint32_t levels = 2;
    for (uint32_t q = 0; q < nq; q++) {
        g2_ctx tmp;  // plan compilation only; no device state
        if (gm_compile(&tmp, &plans[q]) != G2_OK)
            return set_err(c, G2_EINVAL, "plan " + std::to_string(q) + ": " + tmp.err);
        progs[q] = tmp.prog;
        levels = std::max(levels, tmp.prog.n);
    }
    std::vector<uint32_t> order(nq);
    if (cfg->order) {
        std::vector<uint8_t> seen(nq, 0);
        for (uint32_t i = 0; i < nq; i++) {
            const uint32_t o = cfg->order[i];
            if (o >= nq || seen[o]) return set_err(c, G2_EINVAL, "query order is not a permutation");
            seen[o] = 1;
            order[i] = o;
        }
    } else {
        for (uint32_t i = 0; i < nq; i++) order[i] = i;
    }
    const uint64_t* task_off = nullptr;
    uint64_t ntasks = 0;
    int rc;
    if ((rc = batch_task_offsets(c, cfg->window, &task_off, &ntasks))) return rc;
    const uint32_t rpt = std::max<uint32_t>(cfg->roots_per_task, 1);
    const uint64_t per_q = (ntasks + rpt - 1) / rpt;
    const uint32_t wpc = cfg->waves_per_cu ? cfg->waves_per_cu : waves_per_cu(c, WK_GM);
    const uint32_t bw = std::min<uint32_t>(c->tune.block_waves, kGmMaxBlockWaves);
    uint64_t nblocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * wpc / bw);
    const uint64_t cap = std::max<uint64_t>(c->max_degree, 64);
    const uint64_t per_wave = cap * 2 * levels * 4;
    nblocks = fit_arenas(c, nblocks, bw * per_wave);
    if ((rc = ensure(c, c->counters, GCTR_N * 8, "counters"))) return rc;
    if ((rc = ensure(c, c->scratch, nblocks * bw * per_wave, "wave arenas"))) return rc;
    if ((rc = ensure(c, c->batch_progs, (uint64_t)nq * sizeof(GmProgram), "batch programs"))) return rc;
    // per query: order u32, count u64, nodes u64, flag u32
    const uint64_t qbytes = (uint64_t)nq * 24;
    if ((rc = ensure(c, c->batch_q, qbytes, "batch query state"))) return rc;
    auto* ctr = (unsigned long long*)c->counters.ptr;
    auto* qcount = (unsigned long long*)c->batch_q.ptr;
    auto* qn = qcount + nq;
    auto* qord = (uint32_t*)(qn + nq);
    auto* qflag = qord + nq;
    hipError_t e;
    if ((e = hipMemcpyAsync(c->batch_progs.ptr, progs.data(), (uint64_t)nq * sizeof(GmProgram), hipMemcpyHostToDevice,
                            c->stream)))
        return hip_err(c, e, "H2D batch programs");
    if ((e = hipMemcpyAsync(qord, order.data(), (uint64_t)nq * 4, hipMemcpyHostToDevice, c->stream)))
        return hip_err(c, e, "H2D query order");
    if ((e = hipMemsetAsync(ctr, 0, GCTR_N * 8, c->stream))) return hip_err(c, e, "memset counters");
    if ((e = hipMemsetAsync(qcount, 0, (uint64_t)nq * 16, c->stream))) return hip_err(c, e, "memset query counts");
    if ((e = hipMemsetAsync(qflag, 0, (uint64_t)nq * 4, c->stream))) return hip_err(c, e, "memset query flags");
    GmArgs a{};
    a.rp = (const uint64_t*)c->rp.ptr;
    a.cols = (const uint32_t*)c->cols.ptr;
    a.src = (const uint32_t*)c->src.ptr;
    a.task_off = task_off;
    a.ns = c->ns;
    a.ntasks = ntasks;
    a.chunk = cfg->window;
    a.scratch = (uint32_t*)c->scratch.ptr;
    a.cap = cap;
    a.ctr = ctr;
    a.per_root = nullptr;
    GmBatchArgs b{};
    b.progs = (const GmProgram*)c->batch_progs.ptr;
    b.order = qord;
    b.nq = nq;
    b.levels = levels;
    b.rpt = rpt;
    b.per_q = per_q;
    b.qcount = qcount;
    b.nb = NodeBudget{qn, qflag, cfg->node_budget ? cfg->node_budget : ~0ull};
    if ((e = hipEventRecord(c->ev[0], c->stream))) return hip_err(c, e, "event");
    hipLaunchKernelGGL(gm_batch_kernel, dim3(nblocks), dim3(bw * kWave), 0, c->stream, a, b);
    if ((e = hipGetLastError())) return hip_err(c, e, "batch launch");
    if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
    std::vector<uint64_t> hq(2 * (uint64_t)nq);
    std::vector<uint32_t> hf(nq);
    unsigned long long h[GCTR_N];
    if ((e = hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H counters");
    if ((e = hipMemcpyAsync(hq.data(), qcount, (uint64_t)nq * 16, hipMemcpyDeviceToHost, c->stream)))
        return hip_err(c, e, "D2H query counts");
    if ((e = hipMemcpyAsync(hf.data(), qflag, (uint64_t)nq * 4, hipMemcpyDeviceToHost, c->stream)))
        return hip_err(c, e, "D2H query flags");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "batch sync");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    for (uint32_t q = 0; q < nq; q++) {
        if (counts) counts[q] = hq[q];
        if (qnodes) qnodes[q] = hq[nq + q];
        if (over) over[q] = hf[q];
    }
    if (st) {
        st->kernel_ms = ms;
        st->launches = 1;
        st->total_ms = ms;
        st->nodes = h[GCTR_NODES];
        st->max_scratch_words = cap;
    }
    return G2_OK;
}

}  // namespace g2

namespace g2 {

hipError_t gm_launch(const GmArgs& a, const GmProgram& prog, uint64_t blocks, uint32_t block_waves, hipStream_t s) {
    hipLaunchKernelGGL(gm_kernel<false>, dim3(blocks), dim3(block_waves * kWave), 0, s, a, prog);
    return hipGetLastError();
}

}  // namespace g2
