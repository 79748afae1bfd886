// csrc/g2_mce.hip -- maximal clique enumeration on gfx950.
//
// Same search tree as the reference's BK app (app_BK/BK.h:78-474):
//   root node of v:  R = {v}, P = {u in N(v): u > v}, X = {u in N(v): u < v}
//                    (BK.h:323-355; degree-0 roots are skipped, :335)
//   clique test:     P and X empty (BK.h:78-91)
//   pivot:           argmax_{u in P u X} |P & N(u)|, ties to the LAST u in list
//                    order, and list order is vertex-id order (BK.h:282-321)
//   children:        q in P \ N(pivot) in increasing id; earlier q's move to X
//                    (BK.h:445-474 with the Q relabelling of :236-280)
// so totals, size histograms and per-root node counts are bit-exact with it.
//
// MI355X design (nothing of the reference's data layout survives):
//   * one wave64 owns one root at a time (dynamic fetch from a global counter)
//     and walks the root's subtree depth first with an explicit stack, instead
//     of the reference's level-synchronous global queues;
//   * the root's neighbourhood C = N(v) (sorted, position order == id order)
//     becomes bitsets: adjP[k] = N(C[k]) & P0 over the p0 = |P0| later
//     neighbours, and adjX[j] = N(P0[j]) & X0 over the s earlier ones.  Set
//     intersection is then AND + popcount on 64-bit words.  The bitsets are
//     built from the later parts of the members' rows (each edge inside C
//     found at its lower end), short in degeneracy order;
//   * a node is (P, XP) in P0 space (one 64-bit word per lane, lane i holds
//     word i) plus XX in X0 space (a stack frame).  Roots of degree <= 64 keep
//     everything in registers; larger ones keep the rows over X' (the X0
//     members with a P0 neighbour) and the shallow frames in the wave's LDS
//     when they fit, else in the wave's global arena;
//   * pivot scoring gives each candidate to one lane and reduces the key
//     (score << 32 | position) with a 64-bit wave max: max score, then the
//     largest position == the reference's "last maximum" rule.
// Roots whose bitsets exceed the per-wave arena are deferred to a second pass
// with a larger arena (never dropped).
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "g2_device.hpp"
#include "g2_internal.hpp"

namespace g2 {

enum { CTR_NEXT = 0, CTR_CLIQUES, CTR_NODES, CTR_BYTES, CTR_NBIG, CTR_MAXNEED, CTR_ERR, CTR_N };

struct MceArgs {
    const uint64_t* __restrict__ rp;
    const uint32_t* __restrict__ cols;
    const uint32_t* __restrict__ src;
    // task codes: source index | part j << 32 | parts k << 48 (the task walks
    // the root's first-level children c with c mod k == j); null = task t is
    // source t, whole
    const uint64_t* __restrict__ tasks;
    uint64_t first;                      // first task (range runs)
    uint64_t ntasks;                     // one past the last task
    uint64_t* scratch;
    uint64_t scratch_words;              // per wave
    uint64_t* big_list;                  // deferred task codes (null in the large-root pass)
    unsigned long long* ctr;
    unsigned long long* hist;
    uint32_t hist_len;
    unsigned long long* per_root;        // nullable
    unsigned long long* wave_span;       // per wave: first task start, last task end (GPU real-time clock)
    uint32_t small_max;                  // roots with d <= small_max (<= 64) take the register-resident path
    // work donation: a wave that sees idle waves hands its shallowest open
    // frame (its untried children) to a queue the idle waves drain
    unsigned long long* dq;              // DQ_* counters
    uint64_t* pool;                      // record words
    uint64_t pool_words;
    unsigned long long* rec_off;         // per record: offset in pool
    unsigned int* rec_ready;             // per record: published flag
    uint64_t max_records;
    uint32_t donate_every;               // nodes between checks for idle waves (0 = never donate)
};

// donation queue counters: records reserved, records taken, root tasks
// finished (flushed per wave), waves holding a record, waves waiting for one,
// pool words handed out, records published
enum { DQ_RES = 0, DQ_TAKEN, DQ_DONE, DQ_BUSY, DQ_IDLE, DQ_TOP, DQ_MADE, DQ_N };
constexpr uint64_t kDqStride = 16;  // one 128-byte line per counter

// u64 words of the per-root arena: adjP (d*W) + adjX (p0*WX) + (p0+1) frames
__host__ __device__ inline uint64_t mce_need(uint64_t d, uint64_t s) {
    const uint64_t p0 = d - s, W = (p0 + 63) / 64, WX = (s + 63) / 64;
    return d * W + p0 * WX + (p0 + 1) * (3 * W + WX);
}

__device__ inline uint64_t ones(uint64_t nbits) { return nbits >= 64 ? ~0ull : ((1ull << nbits) - 1); }

// The arena is private to the wave, so workgroup-scope ordering is enough to
// make the zero stores precede the atomics and the atomics precede the reads
// (the CU's L1 is shared by the whole workgroup; agent scope would write back
// the XCD's L2 on every root).
__device__ inline void arena_sync() { wave_sync(); }

__device__ inline uint64_t degree_of(const uint64_t* __restrict__ rp, uint32_t u) { return rp[u + 1] - rp[u]; }


// Donation queue.  A record is [source index, absolute depth, P (W words),
// XP (W), Q (W), XX (WX)]: a frame whose untried children Q the taker walks
// exactly as the donor would have (same P/XP progression, same order), so
// the search tree and every count stay the same.
//
// Termination.  A wave adds the number of root tasks it finished to DQ_DONE
// when it leaves the root phase (a wave that never started adds nothing and
// is never waited for), and DQ_BUSY counts waves holding a record.  Records
// are published only by waves inside a root task or holding a record.  An
// idle wave exits once every root task is done, no wave holds a record and
// every reserved record has been taken.  A wave that published records
// claims any still unclaimed before it can exit, so none is left behind.
//
// Coherence without cache maintenance.  A device-scope acquire invalidates
// the XCD's L2 and a device-scope release writes it back, which idle waves
// polling the queue, and donors, would do over and over.  Instead every
// queue word and every record word is accessed with device-scope atomics
// (coherent across XCDs on their own); a writer waits for its earlier
// stores to complete (s_waitcnt) before it issues the store that publishes
// them, and a reader issues a load only after the value it depends on has
// arrived.
__device__ inline unsigned long long ld_dev(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// this wave's earlier memory operations complete before its later ones are
// issued (the signal fences keep the compiler from moving accesses across)
__device__ inline void drain() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// the k lowest set bits of w
__device__ inline uint64_t lowest_bits(uint64_t w, uint32_t k) {
    uint64_t m = 0;
    for (uint32_t i = 0; i < k && w; i++) {
        const uint64_t b = w & (~w + 1);
        m |= b;
        w ^= b;
    }
    return m;
}

// true when idle waves outnumber the records waiting for them
__device__ inline bool hungry(const MceArgs& a) {
    __builtin_amdgcn_wave_barrier();
    uint32_t h = 0;
    if (lane_id() == 0) {
        const uint64_t idle = ld_dev(&a.dq[DQ_IDLE * kDqStride]);
        const uint64_t res = ld_dev(&a.dq[DQ_RES * kDqStride]), taken = ld_dev(&a.dq[DQ_TAKEN * kDqStride]);
        h = idle > res - taken;
    }
    return uniform(h) != 0;
}

// reserve a record of `words` words; false when the pool or the table is full
__device__ inline bool reserve_record(const MceArgs& a, uint64_t words, uint64_t* idx, uint64_t* off) {
    __builtin_amdgcn_wave_barrier();
    unsigned long long i = ~0ull, o = 0;
    if (lane_id() == 0) {
        o = atomicAdd(&a.dq[DQ_TOP * kDqStride], (unsigned long long)words);
        if (o + words <= a.pool_words) {
            unsigned long long c = ld_dev(&a.dq[DQ_RES * kDqStride]);
            while (c < a.max_records) {
                const unsigned long long prev = atomicCAS(&a.dq[DQ_RES * kDqStride], c, c + 1);
                if (prev == c) { i = c; break; }
                c = prev;
            }
        }
    }
    *idx = uniform((uint64_t)i);
    *off = uniform((uint64_t)o);
    return *idx != ~0ull;
}

// the record words were stored (st_dev) by the whole wave: once they have
// completed, flag the record
__device__ inline void publish_record(const MceArgs& a, uint64_t idx, uint64_t off) {
    drain();
    __builtin_amdgcn_wave_barrier();
    if (lane_id() == 0) {
        st_dev(&a.rec_off[idx], off);
        drain();
        __hip_atomic_store(&a.rec_ready[idx], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(&a.dq[DQ_MADE * kDqStride], 1ull);
    }
}

// add (wrapping) to a queue counter after this wave's earlier stores completed
__device__ inline void dq_add(const MceArgs& a, int which, unsigned long long v) {
    drain();
    __builtin_amdgcn_wave_barrier();
    if (lane_id() == 0) atomicAdd(&a.dq[which * kDqStride], v);
}

// Next record for an idle wave (counted in DQ_IDLE): its pool offset, or ~0
// when the run is over.  A claimed record moves the wave from idle to busy.
__device__ inline uint64_t claim_record(const MceArgs& a, uint64_t total_tasks) {
    while (true) {
        __builtin_amdgcn_wave_barrier();
        unsigned long long got = ~0ull - 1;  // nothing yet
        if (lane_id() == 0) {
            const unsigned long long taken = ld_dev(&a.dq[DQ_TAKEN * kDqStride]);
            const unsigned long long res = ld_dev(&a.dq[DQ_RES * kDqStride]);
            if (taken < res) {
                if (atomicCAS(&a.dq[DQ_TAKEN * kDqStride], taken, taken + 1) == taken) {
                    atomicAdd(&a.dq[DQ_BUSY * kDqStride], 1ull);
                    atomicAdd(&a.dq[DQ_IDLE * kDqStride], ~0ull);
                    // the publisher is inside a task, so the flag comes
                    while (__hip_atomic_load(&a.rec_ready[taken], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                        __builtin_amdgcn_s_sleep(2);
                    got = ld_dev(&a.rec_off[taken]);
                }
            } else if (ld_dev(&a.dq[DQ_DONE * kDqStride]) == total_tasks && ld_dev(&a.dq[DQ_BUSY * kDqStride]) == 0 &&
                       ld_dev(&a.dq[DQ_TAKEN * kDqStride]) == ld_dev(&a.dq[DQ_RES * kDqStride])) {
                atomicAdd(&a.dq[DQ_IDLE * kDqStride], ~0ull);
                got = ~0ull;
            }
        }
        got = uniform((uint64_t)got);
        if (got != ~0ull - 1) return got;
        __builtin_amdgcn_s_sleep(127);
    }
}

constexpr int kMaxBlockWaves = 4;  // __launch_bounds__(256)
constexpr int kScoreWords = 2;     // candidate words scored per round (arena path)
constexpr int kLdsArena = 896;     // u64 words of LDS per wave for the walk of a medium root
constexpr int kMinLdsFrames = 16; // frames that must fit next to the rows for the LDS layout
constexpr int kBuildItems = 4;    // neighbourhood-build items per lane in flight (C in LDS)
constexpr uint32_t kPB = 255;      // value buckets over a root's P0 part in the build (C in LDS)

// LDS of one wave for the register-resident path
struct SmallLds {
    uint32_t C[kWave];         // the root's neighbourhood, sorted
    uint32_t pref[kWave];      // exclusive prefix of the streamed row lengths per P0 vertex
    uint64_t start[kWave];     // first cols index of the row slice per P0 vertex
    uint32_t rlen[kWave];      // slice length; bit 31: look C up in the slice instead of streaming it
    uint64_t adjP[kWave];      // N(C[k]) & P0, bit j = P0[j]
    uint64_t adjX[kWave];      // N(P0[j]) & X0, bit k = C[k]
};

__device__ inline uint32_t lds_lower_bound(const uint32_t* A, uint32_t lo, uint32_t hi, uint32_t val) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A[mid] < val) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ inline uint32_t wave_max32(uint32_t x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t y = __shfl_xor(x, m, kWave);
        x = y > x ? y : x;
    }
    return x;
}

// Roots with d <= 64: the whole subtree in registers.  Position k of C maps
// to lane k; lane k keeps row adjP[k] and (k < p0) row adjX[k]; a node is four
// wave-uniform words P, XP (P0 space), XX (X0 space) and Q; the DFS stack is
// spread over the lanes (lane i holds the frame of depth i).  The tree, the
// pivot rule (max |P & N(u)|, ties to the larger position) and the child order
// are those of the arena path, so every count is unchanged.  Clique sizes are
// counted per lane (lane = size, sizes 64 and 65 in hbig at lanes 0 and 1) and
// flushed once per wave.
//
// With rec != nullptr the walk resumes a donated frame (W = WX = 1 here):
// it starts at that node with its untried children, one level above which
// the walk ends.
__device__ inline uint64_t small_walk(const MceArgs& a, uint32_t si, uint32_t s, uint32_t d, uint64_t rowP, uint64_t rowX,
                                      uint32_t part, uint32_t parts, const uint64_t* rec, uint64_t& cliques,
                                      uint64_t& hcount, uint64_t& hbig);

__device__ inline uint64_t small_root(const MceArgs& a, SmallLds& L, uint32_t si, uint32_t v, uint64_t lo, uint32_t d,
                                      uint32_t s, uint32_t part, uint32_t parts, const uint64_t* rec,
                                      uint64_t& cliques, uint64_t& hcount, uint64_t& hbig) {
    const uint32_t lane = lane_id();
    const uint32_t p0 = d - s;
    // ---- neighbourhood bitsets in LDS ----
    if (lane < d) L.C[lane] = a.cols[lo + lane];
    L.adjP[lane] = 0;
    L.adjX[lane] = 0;
    wave_sync();
    // Oriented build: every edge inside C that touches P0 is found in the
    // later part of its lower end's row.  Member k = lane (vertex c) streams
    // N(c) within [v+1, C[d-1]] (c in X0) or [c+1, C[d-1]] (c in P0) and looks
    // each entry up among C's P0 part; a row of at most 64 entries is streamed
    // whole (entries outside the range miss, and a P0 pair found from both
    // ends sets the same bits twice).
    uint32_t len = 0;
    if (lane < d) {
        const uint32_t c = L.C[lane];
        uint64_t e0 = a.rp[c], e1 = a.rp[c + 1];
        if (e1 - e0 > kWave) {
            e0 = lower_bound(a.cols, e0, e1, lane < s ? v + 1 : c + 1);
            e1 = lower_bound(a.cols, e0, e1, L.C[d - 1] + 1);
        }
        L.start[lane] = e0;
        len = (uint32_t)(e1 - e0);
    }
    uint32_t incl = len;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, kWave);
        if (lane >= (uint32_t)o) incl += y;
    }
    L.pref[lane] = incl - len;
    const uint32_t T = read_lane(incl, kWave - 1);
    uint32_t hb_p = 1;  // highest power of two <= p0
    while (hb_p * 2 <= p0) hb_p *= 2;
    wave_sync();
    for (uint32_t f = lane; f < T; f += kWave) {
        // member k whose streamed slice holds entry f (last k with pref[k] <= f)
        uint32_t lo2 = 0, hi2 = d;
        while (hi2 - lo2 > 1) {
            const uint32_t mid = (lo2 + hi2) >> 1;
            if (L.pref[mid] <= f) lo2 = mid; else hi2 = mid;
        }
        const uint32_t k = lo2;
        const uint32_t w = a.cols[L.start[k] + (f - L.pref[k])];
        uint32_t pos = s;  // first P0 position with C >= w, fixed halving steps
        for (uint32_t step = hb_p; step >= 1; step >>= 1)
            if (pos + step <= d && L.C[pos + step - 1] < w) pos += step;
        if (pos < d && L.C[pos] == w && pos != k) {
            const uint32_t j2 = pos - s;
            if (k < s) {  // X0 member k, P0 member j2
                atomicOr((unsigned long long*)&L.adjX[j2], 1ull << k);
                atomicOr((unsigned long long*)&L.adjP[k], 1ull << j2);
            } else {      // P0 members j1, j2
                const uint32_t j1 = k - s;
                atomicOr((unsigned long long*)&L.adjP[s + j1], 1ull << j2);
                atomicOr((unsigned long long*)&L.adjP[s + j2], 1ull << j1);
            }
        }
    }
    wave_sync();
    return small_walk(a, si, s, d, L.adjP[lane], L.adjX[lane], part, parts, rec, cliques, hcount, hbig);
}

// The register walk of a root whose rows are in registers: lane k holds row
// rowP (position k over P0) and, for k < p0, rowX (P0 member k over X).  Also
// used by larger roots whose rows over X' (the X0 members with a neighbour in
// P0) and P0 fit one word each and |X'| + p0 <= 64.
__device__ inline uint64_t small_walk(const MceArgs& a, uint32_t si, uint32_t s, uint32_t d, uint64_t rowP, uint64_t rowX,
                                      uint32_t part, uint32_t parts, const uint64_t* rec, uint64_t& cliques,
                                      uint64_t& hcount, uint64_t& hbig) {
    const uint32_t lane = lane_id();
    const uint32_t p0 = d - s;
    // ---- depth-first walk (depth relative to `base`, the start node's depth) ----
    uint64_t P = ones(p0), XP = 0, XX = ones(s), Q = 0;
    uint32_t base = 0;
    bool fresh = true;
    if (rec) {
        const unsigned long long* r = (const unsigned long long*)rec;
        const uint64_t w = lane < 6 ? ld_dev(r + lane) : 0;
        base = (uint32_t)read_lane(w, 1);
        P = read_lane(w, 2);
        XP = read_lane(w, 3);
        Q = read_lane(w, 4);
        XX = read_lane(w, 5);
        fresh = false;
    }
    uint64_t fP = 0, fXP = 0, fXX = 0, fQ = 0;  // frame of depth == lane
    uint32_t depth = 0, child = 0, since = 0;
    uint64_t nodes = 0;
    while (true) {
        bool descend = false;
        if (fresh) {
            if (base + depth > 0 || part == 0) nodes++;
            // a task that has walked 64 x donate_every nodes is heavy: from then on,
            // every donate_every nodes, look for idle waves
            if (a.donate_every && ++since >= a.donate_every && nodes >= 64ull * a.donate_every) {
                since = 0;
                // hand the upper half of the shallowest frame's untried children
                // to an idle wave (not the root frame of a split root: its
                // children are dealt by part)
                const uint32_t first = (base == 0 && parts > 1) ? 1u : 0u;
                const uint64_t open = ballot(lane < depth && lane >= first && fQ != 0);
                if (open && hungry(a)) {
                    const uint32_t i = __builtin_ctzll(open);
                    uint64_t idx, off;
                    if (reserve_record(a, 6, &idx, &off)) {
                        const uint64_t fp = read_lane(fP, i), fxp = read_lane(fXP, i), fq = read_lane(fQ, i),
                                       fxx = read_lane(fXX, i);
                        // the donor keeps the lower half; the taker's P and XP
                        // already account for those siblings, which come first
                        const uint32_t nq = __popcll(fq);
                        const uint64_t keep = nq >= 2 ? lowest_bits(fq, nq - nq / 2) : 0;
                        const uint64_t w = lane == 0 ? (uint64_t)si : lane == 1 ? (uint64_t)(base + i)
                                         : lane == 2 ? (fp & ~keep) : lane == 3 ? (fxp | keep)
                                         : lane == 4 ? (fq & ~keep) : fxx;
                        if (lane < 6) st_dev((unsigned long long*)&a.pool[off + lane], w);
                        publish_record(a, idx, off);
                        fQ = lane == i ? keep : fQ;
                    }
                }
            }
            if (P == 0) {
                if (XP == 0 && XX == 0) {  // maximal clique, |R| = base + depth + 1
                    cliques++;
                    const uint32_t sz = base + depth + 1;
                    if (sz < kWave) hcount += (lane == sz);
                    else hbig += (lane == sz - kWave);
                }
            } else {
                const uint64_t PX = P | XP;
                const bool cand = lane < s ? ((XX >> lane) & 1) : (lane < d && ((PX >> (lane - s)) & 1));
                const uint32_t key = cand ? ((uint32_t)__popcll(rowP & P) << 8) | lane : 0u;
                const uint32_t pivot = wave_max32(key) & 0xffu;
                Q = P & ~read_lane(rowP, pivot);
            }
            fresh = false;
        }
        if (Q != 0) {
            const uint32_t j = __builtin_ctzll(Q);
            const uint64_t m = 1ull << j;
            Q &= ~m;
            P &= ~m;
            XP |= m;
            if (base + depth == 0 && (child++ % parts) != part) continue;  // another part's child
            const bool mine = lane == depth;
            fP = mine ? P : fP;
            fXP = mine ? XP : fXP;
            fXX = mine ? XX : fXX;
            fQ = mine ? Q : fQ;
            const uint64_t qrow = read_lane(rowP, s + j);
            P &= qrow;
            XP &= qrow;
            XX &= read_lane(rowX, j);
            Q = 0;
            depth++;
            fresh = true;
            descend = true;
        }
        if (!descend) {
            if (depth == 0) break;
            depth--;
            P = read_lane(fP, depth);
            XP = read_lane(fXP, depth);
            XX = read_lane(fXX, depth);
            Q = read_lane(fQ, depth);
        }
    }
    return nodes;
}

template <bool ACCOUNT>
__global__ __launch_bounds__(256, 4) void mce_kernel(MceArgs a) {
    __shared__ unsigned long long lhist[64];
    __shared__ SmallLds small_lds[kMaxBlockWaves];
    __shared__ uint64_t lds_arena[kMaxBlockWaves][kLdsArena];
    for (uint32_t i = threadIdx.x; i < 64; i += blockDim.x) lhist[i] = 0;
    __syncthreads();

    const uint32_t lane = lane_id();
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    SmallLds& SL = small_lds[threadIdx.x / kWave];
    uint64_t* SA = lds_arena[threadIdx.x / kWave];
    uint64_t* ws = a.scratch + gwave * a.scratch_words;
    uint64_t my_cliques = 0, my_nodes = 0, my_bytes = 0;
    uint64_t small_cliques = 0, small_hist = 0, small_big = 0;  // register-resident path tallies
    const uint64_t t_start = wall_clock64();
    uint64_t t_end = t_start;

    // one task: a root (part `part` of `parts`), or a donated frame when rec != nullptr
    auto process = [&](uint64_t code, const uint64_t* rec) {
        const uint32_t si = (uint32_t)code;
        const uint32_t part = (uint32_t)(code >> 32) & 0xffffu;
        const uint32_t parts = (uint32_t)(code >> 48) ? (uint32_t)(code >> 48) : 1u;
        const uint32_t v = uniform(a.src[si]);
        const uint64_t lo = uniform(a.rp[v]);
        const uint64_t d = uniform(a.rp[v + 1]) - lo;
        const uint32_t* __restrict__ C = a.cols + lo;
        uint64_t root_nodes = 0;

        if (d == 0) return;  // no subgraph is generated for an isolated root
        const uint64_t s = uniform(lower_bound(C, 0, d, v));
        const uint64_t p0 = d - s;
        if (p0 == 0) {  // root node with empty P and non-empty X: one node, no clique (never donated)
            if (part == 0) {
                if (ACCOUNT && lane == 0) my_bytes += 5 * (1 + d);
                my_nodes += 1;
                if (a.per_root && lane == 0) atomicAdd(&a.per_root[si], 1ull);
            }
            return;
        }
        if (!ACCOUNT && d <= a.small_max) {
            root_nodes = small_root(a, SL, si, v, lo, (uint32_t)d, (uint32_t)s, part, parts, rec, small_cliques,
                                    small_hist, small_big);
            my_nodes += root_nodes;
            if (a.per_root && lane == 0 && root_nodes) atomicAdd(&a.per_root[si], (unsigned long long)root_nodes);
            t_end = wall_clock64();
            return;
        }
        const uint64_t W = (p0 + 63) / 64, WX = (s + 63) / 64, FW = 3 * W + WX;
        const uint64_t need = mce_need(d, s);
        if (W > kWave) {  // P sets wider than 64 words per lane are not supported
            if (lane == 0) atomicMax(&a.ctr[CTR_ERR], (unsigned long long)need);
            return;
        }
        if (need > a.scratch_words) {
            if (lane == 0) {
                if (a.big_list) {
                    const unsigned long long idx = atomicAdd(&a.ctr[CTR_NBIG], 1ull);
                    a.big_list[idx] = code;
                    atomicMax(&a.ctr[CTR_MAXNEED], (unsigned long long)need);
                } else {
                    atomicMax(&a.ctr[CTR_ERR], (unsigned long long)need);
                }
            }
            return;
        }
        uint64_t* adjP = ws;                 // d rows of W words
        uint64_t* adjX = ws + d * W;         // p0 rows of WX words
        uint64_t* frames = adjX + p0 * WX;   // (p0+1) frames of FW words

        // ---- build the neighbourhood bitsets ------------------------------
        for (uint64_t i = lane; i < d * W + p0 * WX; i += kWave) ws[i] = 0;
        // C is kept in the wave's LDS arena for the build when it fits (the
        // arena is free until the walk's layout is copied in)
        const bool ldsC = d <= 2 * (uint64_t)kLdsArena;
        const uint32_t* Cb = C;
        if (ldsC) {
            uint32_t* Cl = (uint32_t*)SA;
            for (uint64_t i = lane; i < d; i += kWave) Cl[i] = C[i];
            Cb = Cl;
        }
        arena_sync();
        if (ldsC) {
            // Oriented build: every edge inside C that touches P0 is found in
            // the later part of its lower end's row.  Member k (vertex c)
            // streams N(c) within [lo, C[d-1]], lo = v+1 for c in X0 (partners
            // in P0) and c+1 for c in P0 (later partners), and looks each entry
            // up among C's P0 part in LDS.  These later-neighbour parts are
            // short in degeneracy order.  A row of at most 64 entries is
            // streamed whole: entries outside the range miss, and a P0 pair
            // found from both ends sets the same bits twice.
            uint32_t hb_p = 1;  // highest power of two <= p0
            while ((uint64_t)hb_p * 2 <= p0) hb_p *= 2;
            // Value buckets over C's P0 part when it has more than 32
            // entries: bucket(x) = floor((x - P0[0]) * kPB / range) by one high
            // multiply (the identity when P0 spans at most kPB ids), monotone,
            // so a lookup searches one bucket.  pbkt (kPB + 1 u16 in the
            // register path's row area, unused here) holds each bucket's first
            // P0 position.
            const bool pbk = p0 > 32;
            uint16_t* const pbkt = reinterpret_cast<uint16_t*>(SL.adjP);
            uint32_t plo = 0, pmul = 0, pmask = 0, hb_b = 1;
            if (pbk) {
                plo = Cb[s];
                const uint32_t range = Cb[d - 1] - plo + 1;
                pmask = __builtin_amdgcn_readfirstlane(range <= kPB ? ~0u : 0u);
                pmul = __builtin_amdgcn_readfirstlane(range <= kPB ? 0u : (uint32_t)(((uint64_t)kPB << 32) / range));
                auto pb = [&](uint32_t x) { const uint32_t dx = x - plo; return __umulhi(dx, pmul) + (dx & pmask); };
                for (uint32_t i = lane; i <= p0; i += kWave) {  // buckets (pb(P0[i-1]), pb(P0[i])] start at i
                    const uint32_t bi = i < p0 ? pb(Cb[s + i]) : kPB;
                    for (uint32_t b = i == 0 ? 0 : pb(Cb[s + i - 1]) + 1; b <= bi; b++) pbkt[b] = (uint16_t)i;
                }
                wave_sync();
                uint32_t big = 0;
                for (uint32_t b = lane; b < kPB; b += kWave) big = umax32(big, (uint32_t)pbkt[b + 1] - pbkt[b]);
                big = (uint32_t)wave_max((uint64_t)big);
                while (hb_b * 2 <= big) hb_b *= 2;
            }
            for (uint64_t k0 = 0; k0 < d; k0 += kWave) {
                const uint32_t nk = (uint32_t)umin64(kWave, d - k0);
                uint32_t len = 0;
                if (lane < nk) {  // (a row of at most 64 entries is streamed whole)
                    const uint64_t k = k0 + lane;
                    const uint32_t c = Cb[k];
                    uint64_t e0 = a.rp[c], e1 = a.rp[c + 1];
                    if (e1 - e0 > kWave) {
                        e0 = lower_bound(a.cols, e0, e1, k < s ? v + 1 : c + 1);
                        e1 = lower_bound(a.cols, e0, e1, Cb[d - 1] + 1);
                    }
                    SL.start[lane] = e0;
                    len = (uint32_t)(e1 - e0);
                }
                uint32_t incl = len;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint32_t y = __shfl_up(incl, o, kWave);
                    if (lane >= (uint32_t)o) incl += y;
                }
                const uint32_t pref = incl - len;
                SL.pref[lane] = pref;
                const uint32_t T = read_lane(incl, kWave - 1);
                wave_sync();
                // kBuildItems items per lane per round, their loads issued
                // together.  The round's first and last owners come from two
                // ballots (pref is non-decreasing over the lanes); each item
                // searches only between them.
                for (uint32_t f0 = 0; f0 < T; f0 += kBuildItems * kWave) {
                    const uint32_t flast = umin32(f0 + kBuildItems * kWave, T) - 1;
                    const uint32_t o0 = __popcll(ballot(lane < nk && pref <= f0)) - 1;
                    const uint32_t o1 = __popcll(ballot(lane < nk && pref <= flast)) - 1;
                    uint32_t fi[kBuildItems], own[kBuildItems], w[kBuildItems], pos[kBuildItems];
#pragma unroll
                    for (int q = 0; q < kBuildItems; q++) {
                        fi[q] = f0 + q * kWave + lane;
                        own[q] = o0;
                    }
                    if (o1 > o0) {
                        uint32_t hb_o = 1;
                        while (hb_o * 2 <= o1 - o0) hb_o *= 2;
                        for (uint32_t step = hb_o; step >= 1; step >>= 1) {
#pragma unroll
                            for (int q = 0; q < kBuildItems; q++) {
                                const uint32_t c = own[q] + step;
                                const uint32_t pc = SL.pref[umin32(c, o1)];
                                own[q] = (c <= o1 && pc <= fi[q]) ? c : own[q];
                            }
                        }
                    }
                    // (an item past T reads the root's first entry and never hits)
#pragma unroll
                    for (int q = 0; q < kBuildItems; q++)
                        w[q] = a.cols[fi[q] < T ? SL.start[own[q]] + (fi[q] - SL.pref[own[q]]) : lo];
                    uint32_t lim[kBuildItems];
                    uint32_t hb = hb_p;
#pragma unroll
                    for (int q = 0; q < kBuildItems; q++) {
                        pos[q] = (uint32_t)s;
                        lim[q] = (uint32_t)d;
                    }
                    if (pbk) {  // (an entry outside P0's range lands in the last bucket and misses)
                        hb = hb_b;
#pragma unroll
                        for (int q = 0; q < kBuildItems; q++) {
                            const uint32_t dx = w[q] - plo;
                            const uint32_t b = umin32(__umulhi(dx, pmul) + (dx & pmask), kPB - 1);
                            pos[q] = (uint32_t)s + pbkt[b];
                            lim[q] = (uint32_t)s + pbkt[b + 1];
                        }
                    }
                    for (uint32_t step = hb; step >= 1; step >>= 1) {  // first P0 position with C >= w
#pragma unroll
                        for (int q = 0; q < kBuildItems; q++) {
                            const uint32_t c = pos[q] + step;
                            const uint32_t cv = Cb[umin32(c, (uint32_t)d) - 1];
                            pos[q] = (c <= lim[q] && cv < w[q]) ? c : pos[q];
                        }
                    }
#pragma unroll
                    for (int q = 0; q < kBuildItems; q++) {
                        const uint32_t cv = Cb[umin32(pos[q], (uint32_t)d - 1)];
                        if (fi[q] < T && pos[q] < d && cv == w[q]) {
                            const uint64_t k = k0 + own[q], j2 = pos[q] - s;
                            if (k < s) {  // X0 member k, P0 member j2
                                atomicOr((unsigned long long*)&adjX[j2 * WX + (k >> 6)], 1ull << (k & 63));
                                atomicOr((unsigned long long*)&adjP[k * W + (j2 >> 6)], 1ull << (j2 & 63));
                            } else {      // P0 members j1 < j2
                                const uint64_t j1 = k - s;
                                atomicOr((unsigned long long*)&adjP[(s + j1) * W + (j2 >> 6)], 1ull << (j2 & 63));
                                atomicOr((unsigned long long*)&adjP[(s + j2) * W + (j1 >> 6)], 1ull << (j1 & 63));
                            }
                        }
                    }
                }
                wave_sync();  // the chunk's work list is rewritten next
            }
        } else {
            // C in global memory: chunks of 64 later neighbours; within a chunk
            // one flattened work list of (neighbour, item) pairs keeps all lanes
            // busy with independent loads: an item is either one entry of the
            // neighbour's row slice (located in C by binary search) or one
            // position of C (looked up in the slice), whichever is cheaper.
            for (uint64_t j0 = 0; j0 < p0; j0 += kWave) {
                const uint32_t nj = (uint32_t)umin64(kWave, p0 - j0);
                uint32_t len = 0;
                if (lane < nj) {
                    const uint32_t p = C[s + j0 + lane];
                    uint64_t e0 = a.rp[p], e1 = a.rp[p + 1];
                    if (e1 - e0 > kWave) {  // restrict a long row to [C[0], C[d-1]]
                        e0 = lower_bound(a.cols, e0, e1, C[0]);
                        e1 = lower_bound(a.cols, e0, e1, C[d - 1] + 1);
                    }
                    const uint64_t n = e1 - e0;
                    const bool rev = d * (1 + bitlen(n)) < n * (1 + bitlen(d));
                    SL.start[lane] = e0;
                    SL.rlen[lane] = rev ? ((uint32_t)n | 0x80000000u) : (uint32_t)n;
                    len = rev ? (uint32_t)d : (uint32_t)n;
                }
                uint32_t incl = len;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint32_t y = __shfl_up(incl, o, kWave);
                    if (lane >= (uint32_t)o) incl += y;
                }
                SL.pref[lane] = incl - len;
                const uint32_t T = read_lane(incl, kWave - 1);
                wave_sync();
                for (uint32_t f = lane; f < T; f += kWave) {
                    uint32_t lo2 = 0, hi2 = nj;
                    while (hi2 - lo2 > 1) {
                        const uint32_t mid = (lo2 + hi2) >> 1;
                        if (SL.pref[mid] <= f) lo2 = mid; else hi2 = mid;
                    }
                    const uint64_t j = j0 + lo2;
                    const uint32_t rl = SL.rlen[lo2], item = f - SL.pref[lo2];
                    uint64_t k;
                    bool hit;
                    if (rl & 0x80000000u) {
                        k = item;
                        hit = contains(a.cols, SL.start[lo2], SL.start[lo2] + (rl & 0x7fffffffu), C[k]);
                    } else {
                        const uint32_t w = a.cols[SL.start[lo2] + item];
                        k = w < v ? lower_bound(C, 0, s, w) : lower_bound(C, s, d, w);
                        hit = k < d && C[k] == w;
                    }
                    if (hit) {
                        if (k < s) {
                            atomicOr((unsigned long long*)&adjX[j * WX + (k >> 6)], 1ull << (k & 63));
                            atomicOr((unsigned long long*)&adjP[k * W + (j >> 6)], 1ull << (j & 63));
                        } else {
                            const uint64_t kk = k - s;
                            atomicOr((unsigned long long*)&adjP[(s + j) * W + (kk >> 6)], 1ull << (kk & 63));
                        }
                    }
                }
                wave_sync();  // the chunk's work list is rewritten next
            }
        }
        arena_sync();

        // ---- the walk's layout --------------------------------------------
        // X0 members without a neighbour in P0 can neither be the pivot (they
        // score 0, and a P member with a larger position always exists where
        // a pivot is chosen) nor survive into a child's X, so the walk keeps
        // only X' = the others.  When the rows over X' and at least
        // kMinLdsFrames frames fit the wave's LDS, they are copied there; the
        // frames deeper than what fits stay in the global arena.  The tree,
        // pivots and children are unchanged.  (The byte model keeps the
        // reference's full X.)
        uint64_t sE = s, WXE = WX, FWE = FW, DL = 0;
        uint64_t* aP = adjP;
        uint64_t* aX = adjX;
        uint64_t* lfr = frames;
        if (!ACCOUNT && WX <= kWave) {
            uint64_t xmw = 0;  // lane t: X word t of the members with a P0 neighbour
            if (lane < WX)
                for (uint64_t j = 0; j < p0; j++) xmw |= adjX[j * WX + lane];
            const uint32_t cnt = __popcll(xmw);
            uint32_t incl = cnt;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o, kWave);
                if (lane >= (uint32_t)o) incl += y;
            }
            const uint64_t s2 = read_lane(incl, kWave - 1);
            const uint64_t WX2 = (s2 + 63) / 64, FW2 = 3 * W + WX2;
            const uint64_t rows2 = (s2 + p0) * W + p0 * WX2;
            const uint64_t nfr = rows2 < kLdsArena ? (kLdsArena - rows2) / FW2 : 0;
            if (nfr >= kMinLdsFrames) {
                uint64_t* L = SA;
                uint64_t* LX = L + (s2 + p0) * W;
                SL.start[lane] = xmw;  // (the build's work list is no longer needed)
                SL.pref[lane] = incl - cnt;
                for (uint64_t i = lane; i < p0 * WX2; i += kWave) LX[i] = 0;
                for (uint64_t t = 0; t < WX; t++) {  // rows of X' in position order
                    const uint64_t xm = read_lane(xmw, (uint32_t)t);
                    if ((xm >> lane) & 1) {
                        const uint64_t r = read_lane(incl - cnt, (uint32_t)t) + lanes_below(xm), k = t * 64 + lane;
                        for (uint64_t i = 0; i < W; i++) L[r * W + i] = adjP[k * W + i];
                    }
                }
                for (uint64_t i = lane; i < p0 * W; i += kWave) L[s2 * W + i] = adjP[s * W + i];
                wave_sync();
                for (uint64_t f = lane; f < p0 * WX; f += kWave) {  // adjX rows over X' positions
                    const uint64_t j = f / WX, t = f - j * WX;
                    const uint64_t xm = SL.start[t], before = SL.pref[t];
                    uint64_t word = adjX[j * WX + t];
                    while (word) {
                        const uint32_t b = __builtin_ctzll(word);
                        word &= word - 1;
                        const uint64_t c = before + __popcll(xm & ((1ull << b) - 1));
                        atomicOr((unsigned long long*)&LX[j * WX2 + (c >> 6)], 1ull << (c & 63));
                    }
                }
                aP = L;
                aX = LX;
                lfr = LX + p0 * WX2;
                DL = umin64(nfr, p0 + 1);
                sE = s2;
                WXE = WX2;
                FWE = FW2;
                arena_sync();
            }
        }
        // rows over X' and P0 of one word each, |X'| + p0 <= 64: the register walk
        if (!ACCOUNT && DL > 0 && W == 1 && WXE == 1 && sE + p0 <= kWave) {
            const uint64_t rowP = lane < sE + p0 ? aP[lane] : 0;
            const uint64_t rowX = lane < p0 ? aX[lane] : 0;
            root_nodes = small_walk(a, si, (uint32_t)sE, (uint32_t)(sE + p0), rowP, rowX, part, parts, rec, small_cliques,
                                    small_hist, small_big);
            my_nodes += root_nodes;
            if (a.per_root && lane == 0 && root_nodes) atomicAdd(&a.per_root[si], (unsigned long long)root_nodes);
            t_end = wall_clock64();
            return;
        }
        // frame of depth dep: the first DL in LDS, the rest in the global arena
        auto frame = [&](uint64_t dep) -> uint64_t* { return dep < DL ? lfr + dep * FWE : frames + dep * FWE; };

        // ---- depth-first walk of the root's subtree -----------------------
        uint64_t pw = 0, xpw = 0, qw = 0;
        uint64_t base = 0;   // depth of the start node (a donated frame's depth)
        bool fresh = true;
        if (rec) {
            base = uniform((uint64_t)ld_dev((const unsigned long long*)rec + 1));
            uint64_t* F0 = frame(0);
            for (uint64_t i = lane; i < FWE; i += kWave) F0[i] = ld_dev((const unsigned long long*)rec + 2 + i);
            arena_sync();
            if (lane < W) {
                pw = F0[lane];
                xpw = F0[W + lane];
                qw = F0[2 * W + lane];
            }
            fresh = false;
        } else {
            if (lane < W) pw = ones(p0 - (uint64_t)lane * 64);
            uint64_t* F0 = frame(0);
            for (uint64_t i = lane; i < WXE; i += kWave) F0[3 * W + i] = ones(sE - i * 64);
        }
        uint64_t depth = 0;  // relative to base
        uint32_t child = 0;  // first-level children passed so far
        uint32_t since = 0;  // nodes since the last look for idle waves
        while (true) {
            uint64_t* F = frame(depth);
            uint64_t* XX = F + 3 * W;
            bool descend = false;
            // the root node belongs to part 0 of a split root
            const bool own = base + depth > 0 || part == 0;
            if (fresh) {
                if (own) root_nodes++;
                if (!ACCOUNT && a.donate_every && ++since >= a.donate_every &&
                    root_nodes >= 64ull * a.donate_every) {
                    since = 0;
                    if (hungry(a)) {
                        // hand the upper half of the shallowest ancestor frame's untried
                        // children to an idle wave (not the root frame of a split root)
                        for (uint64_t i = (base == 0 && parts > 1) ? 1 : 0; i < depth; i++) {
                            uint64_t* Fi = frame(i);
                            const uint64_t qword = lane < W ? Fi[2 * W + lane] : 0;
                            if (!ballot(qword != 0)) continue;
                            uint64_t idx, off;
                            if (reserve_record(a, 2 + FWE, &idx, &off)) {
                                // the donor keeps the lowest half of Q (over all W words)
                                const uint32_t cnt = __popcll(qword);
                                uint32_t incl = cnt;
#pragma unroll
                                for (int o = 1; o < kWave; o <<= 1) {
                                    const uint32_t y = __shfl_up(incl, o, kWave);
                                    if (lane >= (uint32_t)o) incl += y;
                                }
                                const uint32_t nq = read_lane(incl, kWave - 1), before = incl - cnt;
                                const uint32_t keep_all = nq >= 2 ? nq - nq / 2 : 0;
                                const uint32_t mine = keep_all > before ? min(cnt, keep_all - before) : 0u;
                                const uint64_t keepw = lowest_bits(qword, mine);
                                if (lane < W) {
                                    st_dev((unsigned long long*)&a.pool[off + 2 + lane], Fi[lane] & ~keepw);
                                    st_dev((unsigned long long*)&a.pool[off + 2 + W + lane], Fi[W + lane] | keepw);
                                    st_dev((unsigned long long*)&a.pool[off + 2 + 2 * W + lane], qword & ~keepw);
                                }
                                for (uint64_t k = 3 * W + lane; k < FWE; k += kWave)
                                    st_dev((unsigned long long*)&a.pool[off + 2 + k], Fi[k]);
                                if (lane < 2)
                                    st_dev((unsigned long long*)&a.pool[off + lane], lane == 0 ? (uint64_t)si : base + i);
                                publish_record(a, idx, off);
                                if (lane < W) Fi[2 * W + lane] = keepw;
                            }
                            break;
                        }
                    }
                }
                const bool p_any = ballot(pw != 0) != 0;
                uint64_t xx_or = 0;
                for (uint64_t i = lane; i < WXE; i += kWave) xx_or |= XX[i];
                const bool x_any = (ballot(xpw != 0) | ballot(xx_or != 0)) != 0;
                uint64_t np = 0, nPX = 0;
                if (ACCOUNT) {
                    uint64_t cx = 0;
                    for (uint64_t i = lane; i < WXE; i += kWave) cx += __popcll(XX[i]);
                    np = wave_sum(__popcll(pw));
                    nPX = np + wave_sum(__popcll(xpw)) + wave_sum(cx);
                    if (lane == 0 && own) my_bytes += 5 * (base + depth + 1 + nPX) * (base + depth ? 2 : 1);
                }
                if (!p_any) {
                    if (!x_any && lane == 0) {  // maximal clique R, |R| = base + depth + 1
                        my_cliques++;
                        const uint64_t sz = base + depth + 1;
                        if (sz < 64) atomicAdd(&lhist[sz], 1ull);
                        else atomicAdd(&a.hist[sz < a.hist_len ? sz : a.hist_len - 1], 1ull);
                    }
                } else {
                    // pivot over X0 candidates, then P0 candidates (position order).
                    // Candidate words are taken kScoreWords at a time so the row
                    // loads of their candidates are in flight together.
                    uint64_t best = 0;
                    auto score_words = [&](uint64_t wsrc, uint64_t nz, uint64_t kbase) {
                        while (nz) {
                            uint32_t wq[kScoreWords];
                            uint64_t on[kScoreWords], kq[kScoreWords], sc[kScoreWords];
#pragma unroll
                            for (int q = 0; q < kScoreWords; q++) {
                                wq[q] = nz ? (uint32_t)__builtin_ctzll(nz) : 64u;
                                nz &= nz ? nz - 1 : 0;
                                const uint64_t word = wq[q] < 64 ? read_lane(wsrc, wq[q]) : 0;
                                on[q] = (word >> lane) & 1;
                                kq[q] = kbase + (uint64_t)(wq[q] & 63u) * 64 + lane;
                                sc[q] = 0;
                            }
                            for (uint64_t i = 0; i < W; i++) {
                                const uint64_t pi = read_lane(pw, (uint32_t)i);
#pragma unroll
                                for (int q = 0; q < kScoreWords; q++)
                                    if (on[q]) sc[q] += __popcll(aP[kq[q] * W + i] & pi);
                            }
#pragma unroll
                            for (int q = 0; q < kScoreWords; q++) {
                                if (!on[q]) continue;
                                const uint64_t key = (sc[q] << 32) | kq[q];
                                best = key > best ? key : best;
                                if (ACCOUNT && own) my_bytes += cost_isect(np, degree_of(a.rp, C[kq[q]]));
                            }
                        }
                    };
                    const uint64_t npw = wave_sum(__popcll(pw));  // |P|
                    // column scoring costs ~|P|/4 rounds of loads, candidate scoring
                    // ~(nonzero X words)/kScoreWords rounds: take the cheaper
                    const uint64_t xw0 = (WXE <= kWave && lane < WXE) ? XX[lane] : 0;
                    const uint64_t nzx = __popcll(ballot(xw0 != 0));
                    if (!ACCOUNT && WXE <= kWave && npw < 32 && (npw + 3) / 4 < (nzx + kScoreWords - 1) / kScoreWords) {
                        // Few P members: score all X candidates at once from the
                        // columns of adjX (score(x) = sum over j in P of bit x of
                        // adjX[j], the same number as popc(adjP[x] & P) since the
                        // bitsets are symmetric).  Lane i adds word i of each
                        // member's row into 5 bit-sliced counters, then keeps the
                        // candidates of its word with the largest count.
                        const uint64_t xw = xw0;
                        uint64_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0;
                        for (uint32_t wi = 0; wi < W; wi++) {
                            uint64_t word = read_lane(pw, wi);
                            while (word) {
                                uint64_t r[4];
#pragma unroll
                                for (int q = 0; q < 4; q++) {
                                    const uint64_t j = word ? (uint64_t)wi * 64 + __builtin_ctzll(word) : ~0ull;
                                    word &= word ? word - 1 : 0;
                                    r[q] = (j != ~0ull && lane < WXE) ? aX[j * WXE + lane] : 0;
                                }
#pragma unroll
                                for (int q = 0; q < 4; q++) {
                                    uint64_t c = r[q], t;
                                    t = b0 & c; b0 ^= c; c = t;
                                    t = b1 & c; b1 ^= c; c = t;
                                    t = b2 & c; b2 ^= c; c = t;
                                    t = b3 & c; b3 ^= c; c = t;
                                    b4 ^= c;
                                }
                            }
                        }
                        uint64_t mm = xw, sc = 0;
                        const uint64_t planes[5] = {b0, b1, b2, b3, b4};
#pragma unroll
                        for (int b = 4; b >= 0; b--) {
                            const uint64_t t = mm & planes[b];
                            if (t) {
                                mm = t;
                                sc |= 1ull << b;
                            }
                        }
                        if (xw) {
                            const uint64_t k = (uint64_t)lane * 64 + (63 - __builtin_clzll(mm));
                            const uint64_t key = (sc << 32) | k;
                            best = key > best ? key : best;
                        }
                    } else {
                        for (uint64_t c0 = 0; c0 < WXE; c0 += kWave) {
                            const uint64_t wx = (c0 + lane < WXE) ? XX[c0 + lane] : 0;
                            score_words(wx, ballot(wx != 0), c0 * 64);
                        }
                    }
                    {
                        const uint64_t uw = pw | xpw;
                        score_words(uw, ballot(uw != 0), sE);
                    }
                    best = wave_max(best);
                    const uint64_t pivot = best & 0xffffffffull;
                    const uint64_t* prow = aP + pivot * W;
                    qw = lane < W ? (pw & ~prow[lane]) : 0;
                    if (ACCOUNT && own) {
                        if (lane == 0) my_bytes += cost_isect(np, degree_of(a.rp, C[pivot]));
                        uint64_t nz = ballot(qw != 0);
                        while (nz) {
                            const uint32_t wi = __builtin_ctzll(nz);
                            nz &= nz - 1;
                            const uint64_t word = read_lane(qw, wi);
                            if ((word >> lane) & 1)
                                my_bytes += cost_isect(nPX, degree_of(a.rp, C[s + (uint64_t)wi * 64 + lane]));
                        }
                    }
                }
                fresh = false;
            }
            // next child: lowest q left in Q
            const uint64_t nzq = ballot(qw != 0);
            if (nzq != 0) {
                const uint32_t wi = __builtin_ctzll(nzq);
                const uint64_t word = read_lane(qw, wi);
                const uint32_t b = __builtin_ctzll(word);
                const uint64_t j = (uint64_t)wi * 64 + b;
                if (lane == wi) {
                    const uint64_t m = 1ull << b;
                    qw &= ~m;
                    pw &= ~m;
                    xpw |= m;
                }
                if (base + depth == 0 && (child++ % parts) != part) continue;  // another part's child
                if (lane < W) {
                    F[lane] = pw;
                    F[W + lane] = xpw;
                    F[2 * W + lane] = qw;
                }
                const uint64_t* qrow = aP + (sE + j) * W;
                if (lane < W) {
                    const uint64_t r = qrow[lane];
                    pw &= r;
                    xpw &= r;
                }
                qw = 0;
                const uint64_t* xrow = aX + j * WXE;
                uint64_t* XXc = frame(depth + 1) + 3 * W;
                for (uint64_t i = lane; i < WXE; i += kWave) XXc[i] = XX[i] & xrow[i];
                depth++;
                fresh = true;
                descend = true;
            }
            if (!descend) {
                if (depth == 0) break;
                depth--;
                const uint64_t* Fp = frame(depth);
                if (lane < W) {
                    pw = Fp[lane];
                    xpw = Fp[W + lane];
                    qw = Fp[2 * W + lane];
                }
            }
        }
        my_nodes += root_nodes;
        if (a.per_root && lane == 0 && root_nodes) atomicAdd(&a.per_root[si], (unsigned long long)root_nodes);
    };

    uint64_t roots_done = 0;
    while (true) {
        const uint64_t t = fetch_task(&a.ctr[CTR_NEXT]) + a.first;
        if (t >= a.ntasks) break;
        process(uniform(a.tasks ? a.tasks[t] : (uint64_t)t), nullptr);
        roots_done++;
        t_end = wall_clock64();
    }
    if (a.donate_every) {
        dq_add(a, DQ_IDLE, 1);
        dq_add(a, DQ_DONE, roots_done);
        while (true) {
            const uint64_t off = claim_record(a, a.ntasks - a.first);
            if (off == ~0ull) break;
            const uint64_t* rec = a.pool + off;
            process(uniform((uint64_t)ld_dev((const unsigned long long*)rec)), rec);
            t_end = wall_clock64();
            dq_add(a, DQ_IDLE, 1);
            dq_add(a, DQ_BUSY, ~0ull);
        }
    }
    if (a.wave_span && lane == 0) {
        a.wave_span[2 * gwave] = t_start;
        a.wave_span[2 * gwave + 1] = t_end;
    }

    // ---- flush per-wave and per-block tallies ----------------------------
    if (ACCOUNT) my_bytes = wave_sum(my_bytes);
    if (!ACCOUNT) {
        // small-root tallies: every lane holds the same clique count; lane i
        // holds the cliques of size i
        my_cliques += small_cliques;
        if (small_hist) atomicAdd(&lhist[lane], (unsigned long long)small_hist);
        if (small_big) atomicAdd(&a.hist[umin32(kWave + lane, a.hist_len - 1)], (unsigned long long)small_big);
    }
    if (lane == 0) {
        if (my_cliques) atomicAdd(&a.ctr[CTR_CLIQUES], (unsigned long long)my_cliques);
        if (my_nodes) atomicAdd(&a.ctr[CTR_NODES], (unsigned long long)my_nodes);
        if (ACCOUNT && my_bytes) atomicAdd(&a.ctr[CTR_BYTES], (unsigned long long)my_bytes);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 64; i += blockDim.x)
        if (lhist[i]) atomicAdd(&a.hist[i < a.hist_len ? i : a.hist_len - 1], lhist[i]);
}

// ---------------------------------------------------------------------------
// Batched small roots: 64/G roots per wave, one per group of G lanes (G = 16
// or 32).  Every root here has d <= G, so a node is four G-bit words held by
// all lanes of its group, position k of C maps to lane k of the group, and
// the group's DFS stack sits one frame per lane, as in the register path.
// Group-level exchanges use lane permutes (ds_bpermute) and xor shuffles
// below G, so groups never read each other's lanes.  Same tree, pivot rule
// and child order as the other paths; these roots are never split or donated.
// ---------------------------------------------------------------------------
template <int G>
__device__ inline uint32_t group_read(uint32_t x, uint32_t gbase, uint32_t idx) {
    return (uint32_t)__shfl((int)x, (int)(gbase + idx), kWave);
}

template <int G>
__global__ __launch_bounds__(256) void mce_group_kernel(MceArgs a) {
    static_assert(G == 16 || G == 32, "group width");
    constexpr uint32_t K = kWave / G;
    __shared__ unsigned long long lhist[64];
    __shared__ SmallLds small_lds[kMaxBlockWaves];
    for (uint32_t i = threadIdx.x; i < 64; i += blockDim.x) lhist[i] = 0;
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t gl = lane % G, gbase = lane - gl;
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    SmallLds& L = small_lds[threadIdx.x / kWave];
    uint64_t my_nodes = 0, my_cliques = 0, hcount = 0, hbig = 0;
    const uint64_t t_start = wall_clock64();
    uint64_t t_end = t_start;

    while (true) {
        __builtin_amdgcn_wave_barrier();
        unsigned long long t0 = 0;
        if (lane == 0) t0 = atomicAdd(&a.ctr[CTR_NEXT], (unsigned long long)K);
        t0 = uniform((uint64_t)t0);
        if (t0 >= a.ntasks) break;
        // ---- this group's root ----
        const uint64_t t = t0 + lane / G;
        const bool valid = t < a.ntasks;
        const uint32_t si = valid ? (uint32_t)a.tasks[t] : 0u;
        const uint32_t v = valid ? a.src[si] : 0u;
        const uint64_t lo = valid ? a.rp[v] : 0;
        const uint32_t d = valid ? (uint32_t)(a.rp[v + 1] - lo) : 0u;
        if (gl < d) L.C[lane] = a.cols[lo + gl];
        L.adjP[lane] = 0;
        L.adjX[lane] = 0;
        wave_sync();
        const uint32_t gmask = G == 32 ? 0xffffffffu : 0xffffu;
        const uint32_t smask = (uint32_t)(ballot(gl < d && L.C[lane] < v) >> gbase) & gmask;
        const uint32_t s = __popc(smask), p0 = d - s;
        // oriented build, as in the register path: member gl (vertex c)
        // streams the later part of its row and looks entries up among the
        // group's P0 part; a row of at most 64 entries is streamed whole
        uint32_t len = 0;
        if (gl < d) {
            const uint32_t c = L.C[lane];
            uint64_t e0 = a.rp[c], e1 = a.rp[c + 1];
            if (e1 - e0 > kWave) {
                e0 = lower_bound(a.cols, e0, e1, gl < s ? v + 1 : c + 1);
                e1 = lower_bound(a.cols, e0, e1, L.C[gbase + d - 1] + 1);
            }
            L.start[lane] = e0;
            len = (uint32_t)(e1 - e0);
        }
        uint32_t incl = len;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o, G);
            if (gl >= (uint32_t)o) incl += y;
        }
        L.pref[lane] = incl - len;
        const uint32_t T = group_read<G>(incl, gbase, G - 1);
        wave_sync();
        for (uint32_t f = gl; f < T; f += G) {
            uint32_t lo2 = 0, hi2 = d;
            while (hi2 - lo2 > 1) {
                const uint32_t mid = (lo2 + hi2) >> 1;
                if (L.pref[gbase + mid] <= f) lo2 = mid; else hi2 = mid;
            }
            const uint32_t k = lo2;
            const uint32_t w = a.cols[L.start[gbase + k] + (f - L.pref[gbase + k])];
            const uint32_t pos = lds_lower_bound(L.C + gbase, s, d, w);
            if (pos < d && L.C[gbase + pos] == w && pos != k) {
                const uint32_t j2 = pos - s;
                if (k < s) {  // X0 member k, P0 member j2
                    atomicOr((unsigned long long*)&L.adjX[gbase + j2], 1ull << k);
                    atomicOr((unsigned long long*)&L.adjP[gbase + k], 1ull << j2);
                } else {      // P0 members j1, j2
                    const uint32_t j1 = k - s;
                    atomicOr((unsigned long long*)&L.adjP[gbase + s + j1], 1ull << j2);
                    atomicOr((unsigned long long*)&L.adjP[gbase + s + j2], 1ull << j1);
                }
            }
        }
        wave_sync();
        const uint32_t rowP = (uint32_t)L.adjP[lane];
        const uint32_t rowX = (uint32_t)L.adjX[lane];

        // ---- per-group depth-first walks, stepped together ----
        uint32_t P = p0 >= 32 ? 0xffffffffu : (1u << p0) - 1, XP = 0;
        uint32_t XX = s >= 32 ? 0xffffffffu : (1u << s) - 1, Q = 0;
        uint32_t fP = 0, fXP = 0, fXX = 0, fQ = 0;
        uint32_t depth = 0;
        uint64_t nodes = 0;
        bool fresh = true;
        bool done = d == 0;
        if (!done && p0 == 0) {  // empty P, non-empty X: one node, no clique
            nodes = 1;
            done = true;
        }
        while (ballot(!done)) {
            if (!done) {
                bool descend = false;
                if (fresh) {
                    nodes++;
                    if (P == 0) {
                        if (XP == 0 && XX == 0) {
                            my_cliques += gl == 0;
                            const uint32_t sz = depth + 1;
                            if (sz < G) hcount += (gl == sz);
                            else hbig += (gl == sz - G);
                        }
                    } else {
                        const uint32_t PX = P | XP;
                        const bool cand = gl < s ? ((XX >> gl) & 1) : (gl < d && ((PX >> (gl - s)) & 1));
                        uint32_t key = cand ? ((uint32_t)__popc(rowP & P) << 8) | gl : 0u;
#pragma unroll
                        for (int m = G / 2; m >= 1; m >>= 1) {
                            const uint32_t y = (uint32_t)__shfl_xor((int)key, m, kWave);
                            key = y > key ? y : key;
                        }
                        Q = P & ~group_read<G>(rowP, gbase, key & 0xffu);
                    }
                    fresh = false;
                }
                if (Q != 0) {
                    const uint32_t j = __builtin_ctz(Q);
                    const uint32_t m = 1u << j;
                    Q &= ~m;
                    P &= ~m;
                    XP |= m;
                    const bool mine = gl == depth;
                    fP = mine ? P : fP;
                    fXP = mine ? XP : fXP;
                    fXX = mine ? XX : fXX;
                    fQ = mine ? Q : fQ;
                    const uint32_t qrow = group_read<G>(rowP, gbase, s + j);
                    P &= qrow;
                    XP &= qrow;
                    XX &= group_read<G>(rowX, gbase, j);
                    Q = 0;
                    depth++;
                    fresh = true;
                    descend = true;
                }
                if (!descend) {
                    if (depth == 0) {
                        done = true;
                    } else {
                        depth--;
                        P = group_read<G>(fP, gbase, depth);
                        XP = group_read<G>(fXP, gbase, depth);
                        XX = group_read<G>(fXX, gbase, depth);
                        Q = group_read<G>(fQ, gbase, depth);
                    }
                }
            }
        }
        if (valid && gl == 0) {
            my_nodes += nodes;
            if (a.per_root && nodes) atomicAdd(&a.per_root[si], (unsigned long long)nodes);
        }
        t_end = wall_clock64();
    }
    if (a.wave_span && lane == 0) {
        a.wave_span[2 * gwave] = t_start;
        a.wave_span[2 * gwave + 1] = t_end;
    }
    my_nodes = wave_sum(my_nodes);
    my_cliques = wave_sum(my_cliques);
    if (hcount) atomicAdd(&lhist[gl], (unsigned long long)hcount);
    if (hbig) atomicAdd(&a.hist[umin32(G + gl, a.hist_len - 1)], (unsigned long long)hbig);
    if (lane == 0) {
        if (my_cliques) atomicAdd(&a.ctr[CTR_CLIQUES], (unsigned long long)my_cliques);
        if (my_nodes) atomicAdd(&a.ctr[CTR_NODES], (unsigned long long)my_nodes);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 64; i += blockDim.x)
        if (lhist[i]) atomicAdd(&a.hist[i < a.hist_len ? i : a.hist_len - 1], lhist[i]);
}

// ---------------------------------------------------------------------------
// Task list, built on the device.  Schedule position t holds source
// si = order[t]; it goes to the lane-group list (G > 0, not split, degree
// <= G) or, as its k part codes si | j << 32 | k << 48, to the main list.
// Exclusive scans of the per-position counts keep both lists in schedule
// order.  A full order is also checked to be a permutation of the sources.
// ---------------------------------------------------------------------------
struct TaskBuild {
    const uint32_t* __restrict__ order;  // nullable: identity
    const uint32_t* __restrict__ split;  // nullable: every root whole
    const uint32_t* __restrict__ src;
    const uint32_t* __restrict__ deg;    // degree by vertex id
    uint64_t b, n, ns;                   // positions [b, b + n) of ns sources
    uint32_t G;
    unsigned long long* cnt_main;        // n: part codes of position t
    unsigned int* cnt_small;             // n: 1 when position t is a lane-group root
};

__device__ inline void task_of(const TaskBuild& a, uint64_t i, uint32_t* si, uint32_t* k, bool* small) {
    const uint64_t t = a.b + i;
    *si = a.order ? a.order[t] : (uint32_t)t;
    *k = a.split ? max(1u, min(a.split[*si], 0xffffu)) : 1u;
    *small = a.G && *k == 1 && a.deg[a.src[*si]] <= a.G;
}

__global__ __launch_bounds__(256) void task_count_kernel(TaskBuild a) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t si, k;
        bool small;
        task_of(a, i, &si, &k, &small);
        a.cnt_main[i] = small ? 0 : k;
        a.cnt_small[i] = small ? 1u : 0u;
    }
}

__global__ __launch_bounds__(256) void task_fill_kernel(TaskBuild a, const unsigned long long* off_main,
                                                        const unsigned int* off_small, uint64_t* out, uint64_t main_total) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t si, k;
        bool small;
        task_of(a, i, &si, &k, &small);
        if (small) {
            out[main_total + off_small[i]] = si;
        } else {
            for (uint32_t j = 0; j < k; j++)
                out[off_main[i] + j] = (uint64_t)si | ((uint64_t)j << 32) | ((uint64_t)k << 48);
        }
    }
}

// every entry of order (all ns of them) in range and seen once
__global__ __launch_bounds__(256) void perm_check_kernel(const uint32_t* __restrict__ order, uint64_t ns,
                                                         unsigned int* seen, unsigned int* bad) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = order[i];
        if (o >= ns) {
            atomicOr(bad, 1u);
            continue;
        }
        const unsigned int m = 1u << (o & 31);
        if (atomicOr(&seen[o >> 5], m) & m) atomicOr(bad, 1u);
    }
}

// Build the main and lane-group task lists in c->order (main codes first).
static int build_tasks(g2_ctx* c, const g2_sched* sched, uint32_t G, uint64_t* nmain, uint64_t* nsmall) {
    const bool ranged = sched && sched->end > sched->begin;
    const uint64_t b = ranged ? sched->begin : 0, en = ranged ? sched->end : c->ns, n = en - b;
    const bool has_order = sched && sched->order, has_split = sched && sched->split;
    const uint64_t ns = c->ns;
    const uint64_t seen_words = (ns + 31) / 32;
    if (n > 0x7fffffffull) return set_err(c, G2_EINVAL, "schedules of more than 2^31-1 positions are not supported");
    // layout: order u32[ns] | split u32[ns] | cnt_main u64[n] | off_main u64[n] |
    //         cnt_small u32[n] | off_small u32[n] | seen u32[seen_words] | bad u32 | scan temp
    size_t t1 = 0, t2 = 0;
    hipError_t e;
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, t1, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                               (int)std::max<uint64_t>(n, 1), c->stream)))
        return hip_err(c, e, "scan size");
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, t2, (unsigned int*)nullptr, (unsigned int*)nullptr,
                                               (int)std::max<uint64_t>(n, 1), c->stream)))
        return hip_err(c, e, "scan size");
    const uint64_t temp = std::max<uint64_t>(t1, t2);
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    const uint64_t o_split = al(ns * 4), o_cm = o_split + al(ns * 4), o_om = o_cm + al(n * 8), o_cs = o_om + al(n * 8),
                   o_os = o_cs + al(n * 4), o_seen = o_os + al(n * 4), o_bad = o_seen + al(seen_words * 4),
                   o_tmp = o_bad + 256, total = o_tmp + al(temp);
    int rc;
    if ((rc = ensure(c, c->taskbuild, total, "task-list scratch"))) return rc;
    char* base = (char*)c->taskbuild.ptr;
    auto* d_order = (uint32_t*)base;
    auto* d_split = (uint32_t*)(base + o_split);
    auto* cnt_main = (unsigned long long*)(base + o_cm);
    auto* off_main = (unsigned long long*)(base + o_om);
    auto* cnt_small = (unsigned int*)(base + o_cs);
    auto* off_small = (unsigned int*)(base + o_os);
    auto* seen = (unsigned int*)(base + o_seen);
    auto* bad = (unsigned int*)(base + o_bad);
    if (has_order && (e = hipMemcpyAsync(d_order, sched->order, ns * 4, hipMemcpyHostToDevice, c->stream)))
        return hip_err(c, e, "H2D order");
    if (has_split && (e = hipMemcpyAsync(d_split, sched->split, ns * 4, hipMemcpyHostToDevice, c->stream)))
        return hip_err(c, e, "H2D split");
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((std::max(n, ns) + 255) / 256, 8192));
    if (has_order) {
        if ((e = hipMemsetAsync(seen, 0, seen_words * 4 + 4, c->stream))) return hip_err(c, e, "memset seen");
        if ((e = hipMemsetAsync(bad, 0, 4, c->stream))) return hip_err(c, e, "memset flag");
        hipLaunchKernelGGL(perm_check_kernel, dim3(grid), dim3(256), 0, c->stream, d_order, ns, seen, bad);
    }
    TaskBuild a{has_order ? d_order : nullptr, has_split ? d_split : nullptr, (const uint32_t*)c->src.ptr,
                (const uint32_t*)c->deg.ptr, b, n, ns, G, cnt_main, cnt_small};
    if (n) {
        hipLaunchKernelGGL(task_count_kernel, dim3(grid), dim3(256), 0, c->stream, a);
        size_t tb = temp;
        if ((e = hipcub::DeviceScan::ExclusiveSum(base + o_tmp, tb, cnt_main, off_main, (int)n, c->stream)))
            return hip_err(c, e, "scan");
        tb = temp;
        if ((e = hipcub::DeviceScan::ExclusiveSum(base + o_tmp, tb, cnt_small, off_small, (int)n, c->stream)))
            return hip_err(c, e, "scan");
    }
    unsigned long long tail[2] = {0, 0};
    unsigned int stail[2] = {0, 0}, hbad = 0;
    if (n) {
        if ((e = hipMemcpyAsync(tail, off_main + n - 1, 8, hipMemcpyDeviceToHost, c->stream)) ||
            (e = hipMemcpyAsync(tail + 1, cnt_main + n - 1, 8, hipMemcpyDeviceToHost, c->stream)) ||
            (e = hipMemcpyAsync(stail, off_small + n - 1, 4, hipMemcpyDeviceToHost, c->stream)) ||
            (e = hipMemcpyAsync(stail + 1, cnt_small + n - 1, 4, hipMemcpyDeviceToHost, c->stream)))
            return hip_err(c, e, "D2H task totals");
    }
    if (has_order && (e = hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H flag");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "task list sync");
    if ((e = hipGetLastError())) return hip_err(c, e, "task list kernels");
    if (hbad) return set_err(c, G2_EINVAL, "schedule is not a permutation of the sources");
    *nmain = tail[0] + tail[1];
    *nsmall = (uint64_t)stail[0] + stail[1];
    if ((rc = ensure(c, c->order, std::max<uint64_t>(*nmain + *nsmall, 1) * 8, "task list"))) return rc;
    if (n) hipLaunchKernelGGL(task_fill_kernel, dim3(grid), dim3(256), 0, c->stream, a, off_main, off_small,
                              (uint64_t*)c->order.ptr, *nmain);
    if ((e = hipGetLastError())) return hip_err(c, e, "task fill");
    return G2_OK;
}

// page-locked host arenas of the large-root pass (grow-only, kept with the
// context, not charged to the device budget)
static int host_spill(g2_ctx* c, uint64_t bytes) {
    if (bytes <= c->spill_bytes) return G2_OK;
    if (c->spill) {
        hipError_t e = hipHostFree(c->spill);
        c->spill = nullptr;
        c->spill_bytes = 0;
        if (e != hipSuccess) return hip_err(c, e, "hipHostFree (host spill)");
    }
    hipError_t e = hipHostMalloc(&c->spill, bytes, hipHostMallocNonCoherent);
    if (e != hipSuccess) {
        c->spill = nullptr;
        return hip_err(c, e, "hipHostMalloc (host spill of the large-root arenas)");
    }
    c->spill_bytes = bytes;
    return G2_OK;
}

int mce_run(g2_ctx* c, const g2_sched* sched, uint32_t flags, uint64_t* total, uint64_t* hist,
            uint32_t hist_len, uint64_t* per_root_nodes, g2_stats* st) {
    const bool account = flags & G2_RUN_ACCOUNT;
    const uint32_t hl = hist ? hist_len : 2;
    const uint32_t bw = std::min<uint32_t>(c->tune.block_waves, kMaxBlockWaves);
    uint64_t nblocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * waves_per_cu(c, WK_MCE) / bw);
    const uint64_t sw = c->tune.wave_scratch_words;
    // under a budget the grid shrinks until its wave arenas fit what the
    // budget leaves (fewer resident waves, same result)
    if (c->budget) {
        const uint64_t other = c->used - c->scratch.bytes;
        const uint64_t room = c->budget > other ? c->budget - other : 0;
        while (nblocks > 1 && nblocks * bw * sw * 8 > room) nblocks /= 2;
    }
    const uint64_t nwaves = nblocks * bw;
    int rc;
    if ((rc = ensure(c, c->counters, CTR_N * 8, "counters"))) return rc;
    if ((rc = ensure(c, c->hist, (uint64_t)hl * 8, "histogram"))) return rc;
    if ((rc = ensure(c, c->scratch, nwaves * sw * 8, "wave arenas"))) return rc;
    if ((rc = ensure(c, c->spans, nwaves * 16, "wave spans"))) return rc;
    if (per_root_nodes && (rc = ensure(c, c->per_root, std::max<uint64_t>(c->ns, 1) * 8, "per-root counts"))) return rc;
    // task list: sources in schedule order, heavy ones dealt into parts
    const uint64_t* tasks = nullptr;
    uint64_t first = 0, ntasks = c->ns;
    // batched small roots: roots of degree <= G go to the lane-group kernel
    const uint32_t rpw = c->tune.mce_roots_per_wave;
    const uint32_t G = account ? 0u : rpw >= 4 ? 16u : rpw >= 2 ? 32u : 0u;
    uint64_t nsmall = 0;  // roots for the lane-group kernel, after the main list
    if (G || (sched && (sched->order || sched->split))) {
        uint64_t nmain = 0;
        if ((rc = build_tasks(c, sched, G, &nmain, &nsmall))) return rc;
        tasks = (const uint64_t*)c->order.ptr;
        ntasks = nmain;
    } else if (sched && sched->end > sched->begin) {
        // a range of the source order needs no task list: tasks [begin, end)
        first = sched->begin;
        ntasks = sched->end;
    }
    // donation queue: counters, record table and a word pool (a record is a frame plus two words)
    const uint32_t every = account ? 0u : (c->tune.mce_donate_every == 0xffffffffu ? 0u : c->tune.mce_donate_every);
    const uint64_t max_records = 1ull << 16, pool_words = 1ull << 23;
    const uint64_t dq_words = DQ_N * kDqStride;
    if ((rc = ensure(c, c->donate, dq_words * 8 + max_records * 12 + pool_words * 8, "donation queue"))) return rc;
    // the one buffer that grows with the number of root tasks comes last, so
    // that a smaller range of roots fits where the whole view did not
    if ((rc = ensure(c, c->list, std::max<uint64_t>(ntasks - first, 1) * 8, "deferred roots"))) return rc;
    auto* dq = (unsigned long long*)c->donate.ptr;
    auto* rec_off = dq + dq_words;
    auto* rec_ready = (unsigned int*)(rec_off + max_records);
    auto* pool = (uint64_t*)(rec_ready + max_records);
    auto reset_queue = [&](uint64_t waves) -> hipError_t {
        (void)waves;
        hipError_t err = hipMemsetAsync(dq, 0, dq_words * 8, c->stream);
        if (err == hipSuccess) err = hipMemsetAsync(rec_ready, 0, max_records * 4, c->stream);
        return err;
    };
    hipError_t e;
    auto* ctr = (unsigned long long*)c->counters.ptr;
    if ((e = hipMemsetAsync(ctr, 0, CTR_N * 8, c->stream))) return hip_err(c, e, "memset counters");
    if ((e = hipMemsetAsync(c->hist.ptr, 0, (uint64_t)hl * 8, c->stream))) return hip_err(c, e, "memset hist");
    if (per_root_nodes && (e = hipMemsetAsync(c->per_root.ptr, 0, c->ns * 8, c->stream))) return hip_err(c, e, "memset per-root");

    MceArgs a{};
    a.rp = (const uint64_t*)c->rp.ptr;
    a.cols = (const uint32_t*)c->cols.ptr;
    a.src = (const uint32_t*)c->src.ptr;
    a.tasks = tasks;
    a.first = first;
    a.ntasks = ntasks;
    a.scratch = (uint64_t*)c->scratch.ptr;
    a.scratch_words = sw;
    a.big_list = (uint64_t*)c->list.ptr;
    a.ctr = ctr;
    a.hist = (unsigned long long*)c->hist.ptr;
    a.hist_len = hl;
    a.per_root = per_root_nodes ? (unsigned long long*)c->per_root.ptr : nullptr;
    a.wave_span = (unsigned long long*)c->spans.ptr;
    a.small_max = std::min<uint32_t>(c->tune.mce_small_max, kWave);
    a.dq = dq;
    a.pool = pool;
    a.pool_words = pool_words;
    a.rec_off = rec_off;
    a.rec_ready = rec_ready;
    a.max_records = max_records;
    a.donate_every = every;

    auto launch_shape = [&](uint64_t blocks, uint32_t waves) {
        if (account) hipLaunchKernelGGL(mce_kernel<true>, dim3(blocks), dim3(waves * kWave), 0, c->stream, a);
        else hipLaunchKernelGGL(mce_kernel<false>, dim3(blocks), dim3(waves * kWave), 0, c->stream, a);
        return hipGetLastError();
    };
    auto launch = [&](uint64_t blocks) { return launch_shape(blocks, bw); };
    uint64_t big_blocks = 0;
    uint32_t big_bw = bw;
    float k1 = 0, k2 = 0, tot = 0;
    uint32_t launches = 1;
    if ((e = reset_queue(nwaves))) return hip_err(c, e, "donation queue reset");
    if ((e = hipEventRecord(c->ev[0], c->stream))) return hip_err(c, e, "event");
    if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
    if ((e = launch(nblocks))) return hip_err(c, e, "mce launch");
    if ((e = hipEventRecord(c->ev[2], c->stream))) return hip_err(c, e, "event");
    unsigned long long h_ctr[CTR_N];
    std::vector<unsigned long long> spans(2 * nwaves);
    if ((e = hipMemcpyAsync(h_ctr, ctr, sizeof(h_ctr), hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H counters");
    if ((e = hipMemcpyAsync(spans.data(), c->spans.ptr, nwaves * 16, hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H spans");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "mce sync");
    (void)hipEventElapsedTime(&k1, c->ev[1], c->ev[2]);
    double idle = 0;
    {
        unsigned long long lo = ~0ull, hi = 0;
        double busy = 0;
        for (uint64_t w = 0; w < nwaves; w++) {
            lo = std::min(lo, spans[2 * w]);
            hi = std::max(hi, spans[2 * w + 1]);
            busy += (double)(spans[2 * w + 1] - spans[2 * w]);
        }
        if (hi > lo) idle = 1.0 - busy / ((double)nwaves * (double)(hi - lo));
    }
    unsigned long long h_made = 0;
    if ((e = hipMemcpy(&h_made, dq + DQ_MADE * kDqStride, 8, hipMemcpyDeviceToHost)))
        return hip_err(c, e, "D2H donation counters");
    uint64_t donations = h_made;
    const uint64_t nbig = h_ctr[CTR_NBIG];
    uint64_t spill_bytes = 0;
    if (nbig) {
        // large-root pass: arenas sized for the largest deferred root, as
        // many waves as the budget holds (whole workgroups of bw waves, or
        // one smaller workgroup)
        const uint64_t need = h_ctr[CTR_MAXNEED];
        uint64_t bwaves = std::min<uint64_t>(c->tune.big_waves, nbig);
        while (bwaves > 1 && c->budget && c->used - c->big_scratch.bytes + bwaves * need * 8 > c->budget) bwaves /= 2;
        big_bw = (uint32_t)std::min<uint64_t>(bw, bwaves);
        bwaves = bwaves / big_bw * big_bw;
        uint64_t* arenas = nullptr;
        if ((rc = ensure(c, c->big_scratch, bwaves * need * 8, "large-root arenas")) == G2_OK) {
            arenas = (uint64_t*)c->big_scratch.ptr;
        } else if (rc == G2_EBUDGET) {
            // Host spill: the device budget cannot hold even one wave's arena,
            // so the arenas go to page-locked host memory that the kernel
            // reads and writes over PCIe (coarse-grained: each arena is used
            // by one wave only, its atomics stay in that XCD's L2).  Slow, and
            // only for budgets far below the roots' needs; the result is the
            // same.  The reference's host spill is dumpToHost / loadFromHost
            // (app_BK/BK.h:357-398, system/subgraph_container.h:213-216).
            bwaves = std::min<uint64_t>(std::max<uint32_t>(c->tune.big_waves, 1), nbig);
            big_bw = (uint32_t)std::min<uint64_t>(bw, bwaves);
            bwaves = bwaves / big_bw * big_bw;
            if ((rc = host_spill(c, bwaves * need * 8))) return rc;
            arenas = (uint64_t*)c->spill;
            spill_bytes = bwaves * need * 8;
        } else {
            return rc;
        }
        big_blocks = bwaves / big_bw;
        a.tasks = (const uint64_t*)c->list.ptr;
        a.first = 0;
        a.ntasks = nbig;
        a.scratch = arenas;
        a.scratch_words = need;
        a.big_list = nullptr;
        a.wave_span = nullptr;
        if ((e = hipMemsetAsync(ctr + CTR_NEXT, 0, 8, c->stream))) return hip_err(c, e, "memset next");
        if ((e = reset_queue(bwaves))) return hip_err(c, e, "donation queue reset");
        if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
        if ((e = launch_shape(big_blocks, big_bw))) return hip_err(c, e, "mce large-root launch");
        if ((e = hipEventRecord(c->ev[2], c->stream))) return hip_err(c, e, "event");
        launches = 2;
    }
    float k3 = 0;
    bool k2_done = false;
    if (nsmall) {
        if (nbig) {  // the large-root pass's events are reused below
            if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "mce sync");
            (void)hipEventElapsedTime(&k2, c->ev[1], c->ev[2]);
            k2_done = true;
        }
        // lane-group kernel over the small roots
        a.tasks = tasks + ntasks;
        a.first = 0;
        a.ntasks = nsmall;
        a.wave_span = nullptr;
        a.donate_every = 0;
        const uint64_t gblocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * (c->tune.waves_per_cu ? c->tune.waves_per_cu : 32) / bw);
        if ((e = hipMemsetAsync(ctr + CTR_NEXT, 0, 8, c->stream))) return hip_err(c, e, "memset next");
        if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
        if (G == 16) hipLaunchKernelGGL(mce_group_kernel<16>, dim3(gblocks), dim3(bw * kWave), 0, c->stream, a);
        else hipLaunchKernelGGL(mce_group_kernel<32>, dim3(gblocks), dim3(bw * kWave), 0, c->stream, a);
        if ((e = hipGetLastError())) return hip_err(c, e, "mce lane-group launch");
        if ((e = hipEventRecord(c->ev[2], c->stream))) return hip_err(c, e, "event");
        if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "mce lane-group sync");
        (void)hipEventElapsedTime(&k3, c->ev[1], c->ev[2]);
        launches++;
    }
    if ((e = hipEventRecord(c->ev[3], c->stream))) return hip_err(c, e, "event");
    if ((e = hipMemcpyAsync(h_ctr, ctr, sizeof(h_ctr), hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H counters");
    if (hist && (e = hipMemcpyAsync(hist, c->hist.ptr, (uint64_t)hl * 8, hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H hist");
    if (per_root_nodes && c->ns &&
        (e = hipMemcpyAsync(per_root_nodes, c->per_root.ptr, c->ns * 8, hipMemcpyDeviceToHost, c->stream)))
        return hip_err(c, e, "D2H per-root");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "mce sync");
    if (nbig) {
        if (!k2_done) (void)hipEventElapsedTime(&k2, c->ev[1], c->ev[2]);
        if ((e = hipMemcpy(&h_made, dq + DQ_MADE * kDqStride, 8, hipMemcpyDeviceToHost)))
            return hip_err(c, e, "D2H donation counters");
        donations += h_made;
    }
    (void)hipEventElapsedTime(&tot, c->ev[0], c->ev[3]);
    if (h_ctr[CTR_ERR])
        return set_err(c, G2_EBUDGET, "a root needs more than 4096 later neighbours or did not fit the large-root arena");
    if (total) *total = h_ctr[CTR_CLIQUES];
    if (st) {
        st->kernel_ms = (double)k1 + (double)k2 + (double)k3;
        st->launches = launches;
        st->total_ms = tot;
        st->nodes = h_ctr[CTR_NODES];
        st->alg_bytes = account ? h_ctr[CTR_BYTES] : 0;
        st->big_roots = nbig;
        st->max_scratch_words = h_ctr[CTR_MAXNEED];
        st->wave_idle_frac = idle;
        st->donations = donations;
        st->host_spill_bytes = spill_bytes;
    }
    return G2_OK;
}

}  // namespace g2
