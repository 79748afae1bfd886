// csrc/g2_gm_local.hip -- oriented local-bitmap matching for root-dominated plans.
//
// Applies when the plan's root (position 0) is a back neighbour of, and
// ordered below (M(u_i) > M(u_0)), every other position -- every k-clique
// plan among the reference's presets -- and the run covers every vertex as a
// root.  Then each match lies in U(v0) = { w in N(v0) : w above v0 } and the
// plan's back-neighbour / condition tables become bit operations over U.
//
// Orientation: "above" compares the key (degree, id) instead of the id.  The
// plan's condition pairs are a complete symmetry breaking (checked for every
// ranking of the query vertices when the plan is set), so for any total order
// of the data vertices exactly one member of each automorphism class of an
// embedding satisfies them: the total count over all roots is the same number
// the reference's id order gives (app_gmatch/gmatch.h:559-588), while |U| is
// bounded by O(sqrt(E)) instead of by the largest degree.
//
// Per root (one wave64 each, dynamic fetch):
//   1. U = ballot-compacted { w in N(v0) : key(w) > key(v0) }, id order;
//   2. rank of each member inside U by key (local bit index = rank), so the
//      order conditions become contiguous bit ranges;
//   3. rows[r] = N(U_r) & U as m-bit sets, one wave pass per member (stream
//      N(u) against U, or look U up in N(u) when that row is long);
//   4. each lane walks the matches of one position-1 candidate depth first:
//      cand(i) = AND of the rows of its back neighbours & order ranges &
//      NON_EQUAL holes; the last position is a popcount.
// Roots whose arena need exceeds the per-wave size are deferred to a second
// launch with larger arenas.
#include <algorithm>
#include <vector>

#include "g2_device.hpp"
#include "g2_internal.hpp"

namespace g2 {

enum { LCTR_NEXT = 0, LCTR_COUNT, LCTR_NODES, LCTR_NBIG, LCTR_MAXNEED, LCTR_ERR, LCTR_N };

struct LocalArgs {
    const uint64_t* __restrict__ rp;
    const uint32_t* __restrict__ cols;
    const uint32_t* __restrict__ deg;
    const uint32_t* __restrict__ list;  // roots of the large-root pass (nullable: roots = [0, ntasks))
    uint64_t ntasks;
    LocalProgram prog;
    uint64_t* scratch;
    uint64_t scratch_words;
    uint32_t* big_list;                 // deferred roots (null in the large-root pass)
    unsigned long long* ctr;
};

// u64 words of a root's arena: ids (m u32) + keys (m) + ranks (m u32) +
// rows (m*W) + per-lane DFS stacks (64 lanes * (n-2) levels * W)
__host__ __device__ inline uint64_t local_need(uint64_t m, uint32_t n) {
    const uint64_t W = (m + 63) / 64;
    const uint64_t lv = n > 2 ? n - 2 : 0;
    return (m + 1) / 2 + m + (m + 1) / 2 + m * W + 64 * lv * W + 8;
}

__device__ inline uint64_t bits_above(uint64_t r, uint64_t w) {  // bits > r within word w
    const uint64_t lo = w * 64;
    if (r + 1 <= lo) return ~0ull;
    if (r >= lo + 63) return 0;
    return ~0ull << (r - lo + 1);
}

__device__ inline uint64_t bits_below(uint64_t r, uint64_t w) {  // bits < r within word w
    const uint64_t lo = w * 64;
    if (r <= lo) return 0;
    if (r >= lo + 64) return ~0ull;
    return (1ull << (r - lo)) - 1;
}

// candidates of position i as a W-word set given the matched ranks M[1..i-1]
__device__ inline uint64_t cand_word(const LocalProgram& P, uint32_t i, const uint32_t* M,
                                     const uint64_t* rows, uint64_t W, uint64_t m, uint64_t w) {
    uint64_t x = (w + 1) * 64 <= m ? ~0ull : ((1ull << (m - w * 64)) - 1);
    uint32_t b = P.bn_mask[i];
    while (b) {
        const uint32_t j = __builtin_ctz(b);
        b &= b - 1;
        x &= rows[(uint64_t)M[j] * W + w];
    }
    uint32_t g = P.gt_mask[i];
    while (g) {
        const uint32_t j = __builtin_ctz(g);
        g &= g - 1;
        x &= bits_above(M[j], w);
    }
    uint32_t l = P.lt_mask[i];
    while (l) {
        const uint32_t j = __builtin_ctz(l);
        l &= l - 1;
        x &= bits_below(M[j], w);
    }
    uint32_t q = P.ne_mask[i];
    while (q) {
        const uint32_t j = __builtin_ctz(q);
        q &= q - 1;
        if ((M[j] >> 6) == w) x &= ~(1ull << (M[j] & 63));
    }
    return x;
}

template <int N>
__global__ __launch_bounds__(256) void gm_local_kernel(LocalArgs a) {
    const uint32_t lane = lane_id();
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    uint64_t* ws = a.scratch + gwave * a.scratch_words;
    uint64_t lane_total = 0, nodes = 0;
    const LocalProgram& P = a.prog;

    while (true) {
        const uint64_t t = fetch_task(&a.ctr[LCTR_NEXT]);
        if (t >= a.ntasks) break;
        const uint32_t v0 = uniform(a.list ? a.list[t] : (uint32_t)t);
        const uint64_t lo = uniform(a.rp[v0]), hi = uniform(a.rp[v0 + 1]);
        if (hi - lo < (uint64_t)(N - 1)) continue;
        const uint64_t key0 = ((uint64_t)a.deg[v0] << 32) | v0;

        // 1. U in id order; the arena is only touched once the need is known,
        //    so count first
        uint64_t m = 0;
        for (uint64_t base = lo; base < hi; base += kWave) {
            const uint64_t e = base + lane;
            bool up = false;
            if (e < hi) {
                const uint32_t w = a.cols[e];
                up = (((uint64_t)a.deg[w] << 32) | w) > key0;
            }
            m += __popcll(ballot(up));
        }
        if (m < (uint64_t)(N - 1)) continue;
        const uint64_t W = (m + 63) / 64;
        const uint64_t need = local_need(m, N);
        if (need > a.scratch_words) {
            if (lane == 0) {
                if (a.big_list) {
                    const unsigned long long idx = atomicAdd(&a.ctr[LCTR_NBIG], 1ull);
                    a.big_list[idx] = v0;
                    atomicMax(&a.ctr[LCTR_MAXNEED], (unsigned long long)need);
                } else {
                    atomicMax(&a.ctr[LCTR_ERR], (unsigned long long)need);
                }
            }
            continue;
        }
        uint32_t* Uid = (uint32_t*)ws;
        uint64_t* Ukey = ws + (m + 1) / 2;
        uint32_t* Urank = (uint32_t*)(Ukey + m);
        uint64_t* rows = Ukey + m + (m + 1) / 2;
        uint64_t* stacks = rows + m * W;
        uint64_t k = 0;
        for (uint64_t base = lo; base < hi; base += kWave) {
            const uint64_t e = base + lane;
            bool up = false;
            uint32_t w = 0;
            uint64_t kw = 0;
            if (e < hi) {
                w = a.cols[e];
                kw = ((uint64_t)a.deg[w] << 32) | w;
                up = kw > key0;
            }
            const uint64_t bm = ballot(up);
            if (up) {
                const uint64_t p = k + lanes_below(bm);
                Uid[p] = w;
                Ukey[p] = kw;
            }
            k += __popcll(bm);
        }
        for (uint64_t i = lane; i < m * W; i += kWave) rows[i] = 0;
        wave_sync();
        // 2. rank of every member by key
        for (uint64_t j = lane; j < m; j += kWave) {
            const uint64_t kj = Ukey[j];
            uint32_t r = 0;
            for (uint64_t i = 0; i < m; i++) r += Ukey[i] < kj;
            Urank[j] = r;
        }
        wave_sync();
        // 3. rows over ranks
        const uint32_t umin = Uid[0], umax = Uid[m - 1];
        for (uint64_t j = 0; j < m; j++) {
            const uint32_t u = Uid[j];
            const uint64_t rj = Urank[j];
            const uint64_t ulo = a.rp[u], uhi = a.rp[u + 1];
            const uint64_t s0 = lower_bound(a.cols, ulo, uhi, umin);
            const uint64_t s1 = lower_bound(a.cols, s0, uhi, umax + 1);
            uint64_t* row = rows + rj * W;
            if (s1 - s0 <= m * (1 + bitlen(m))) {
                for (uint64_t e = s0 + lane; e < s1; e += kWave) {
                    const uint32_t w = a.cols[e];
                    const uint64_t p = lower_bound(Uid, 0, m, w);
                    if (p < m && Uid[p] == w) {
                        const uint32_t rp_ = Urank[p];
                        atomicOr((unsigned long long*)&row[rp_ >> 6], 1ull << (rp_ & 63));
                    }
                }
            } else {
                for (uint64_t p = lane; p < m; p += kWave) {
                    if (contains(a.cols, s0, s1, Uid[p])) {
                        const uint32_t rp_ = Urank[p];
                        atomicOr((unsigned long long*)&row[rp_ >> 6], 1ull << (rp_ & 63));
                    }
                }
            }
        }
        wave_sync();
        // 4. per-lane depth-first walk from each position-1 candidate
        if (N == 2) {
            if (lane == 0) lane_total += m;
            continue;
        }
        uint64_t* stk = stacks + (uint64_t)lane * (N - 2) * W;  // level i at stk[(i-2)*W]
        uint32_t M[N];
        uint64_t cur[N];
        for (uint64_t r1 = lane; r1 < m; r1 += kWave) {
            M[1] = (uint32_t)r1;
            // level 2
            if (N == 3) {
                for (uint64_t w = 0; w < W; w++) lane_total += __popcll(cand_word(P, 2, M, rows, W, m, w));
                nodes++;
                continue;
            }
            for (uint64_t w = 0; w < W; w++) stk[w] = cand_word(P, 2, M, rows, W, m, w);
            nodes++;
            uint32_t level = 2;
            cur[2] = 0;
            while (true) {
                uint64_t* cs = stk + (uint64_t)(level - 2) * W;
                // next member of cand(level) at or after cur[level]
                uint64_t pos = cur[level];
                bool found = false;
                while (pos < m) {
                    const uint64_t w = pos >> 6;
                    const uint64_t x = cs[w] & (~0ull << (pos & 63));
                    if (x) {
                        pos = w * 64 + __builtin_ctzll(x);
                        found = true;
                        break;
                    }
                    pos = (w + 1) * 64;
                }
                if (!found) {
                    if (level == 2) break;
                    level--;
                    continue;
                }
                cur[level] = pos + 1;
                M[level] = (uint32_t)pos;
                const uint32_t nx = level + 1;
                nodes++;
                if (nx == N - 1) {
                    for (uint64_t w = 0; w < W; w++) lane_total += __popcll(cand_word(P, nx, M, rows, W, m, w));
                } else {
                    uint64_t* ns = stk + (uint64_t)(nx - 2) * W;
                    for (uint64_t w = 0; w < W; w++) ns[w] = cand_word(P, nx, M, rows, W, m, w);
                    cur[nx] = 0;
                    level = nx;
                }
            }
        }
    }
    const uint64_t tot = wave_sum(lane_total);
    const uint64_t nd = wave_sum(nodes);
    if (lane == 0) {
        if (tot) atomicAdd(&a.ctr[LCTR_COUNT], (unsigned long long)tot);
        if (nd) atomicAdd(&a.ctr[LCTR_NODES], (unsigned long long)nd);
    }
}

// ---- plan analysis (host) ---------------------------------------------------

// The plan's order conditions select exactly one member of every
// automorphism class for every ranking of the query vertices.
static bool complete_symmetry_breaking(uint32_t n, const uint32_t adj[G2_MAXQ],
                                       const std::vector<std::pair<uint32_t, uint32_t>>& lt) {
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    std::vector<std::vector<uint32_t>> aut;
    do {
        bool ok = true;
        for (uint32_t x = 0; x < n && ok; x++)
            for (uint32_t y = 0; y < n && ok; y++)
                if (((adj[x] >> y) & 1) != ((adj[perm[x]] >> perm[y]) & 1)) ok = false;
        if (ok) aut.push_back(perm);
    } while (std::next_permutation(perm.begin(), perm.end()));
    // the check visits n! rankings x |Aut| automorphisms; beyond ~1e8 steps
    // the plan simply keeps the canonical path
    uint64_t fact = 1;
    for (uint32_t i = 2; i <= n; i++) fact *= i;
    if (fact * aut.size() > 100000000ull) return false;
    std::vector<uint32_t> rank(n);
    for (uint32_t i = 0; i < n; i++) rank[i] = i;
    do {  // rank[v] = order of the data vertex matched to query vertex v
        uint32_t hits = 0;
        for (const auto& s : aut) {
            bool ok = true;
            for (const auto& c : lt)  // M(a) < M(b) evaluated on the relabelled match
                if (rank[s[c.first]] >= rank[s[c.second]]) { ok = false; break; }
            hits += ok;
            if (hits > 1) return false;
        }
        if (hits != 1) return false;
    } while (std::next_permutation(rank.begin(), rank.end()));
    return true;
}

void gm_local_compile(g2_ctx* c, const g2_plan* p) {
    c->local_ok = false;
    const uint32_t n = p->n;
    LocalProgram L{};
    L.n = n;
    uint32_t adj[G2_MAXQ] = {0};  // query adjacency over positions
    std::vector<std::pair<uint32_t, uint32_t>> lt;  // (a,b): M(a) < M(b), positions
    for (uint32_t i = 1; i < n; i++) {
        const GmStep& s = c->prog.step[i];
        bool root_bn = false, root_gt = false;
        for (uint32_t k = 0; k < s.nbn; k++) {
            const uint32_t b = s.bn[k];
            adj[i] |= 1u << b;
            adj[b] |= 1u << i;
            if (b == 0) root_bn = true; else L.bn_mask[i] |= 1u << b;
        }
        for (uint32_t k = 0; k < s.ncond; k++) {
            const uint32_t j = s.cpos[k], op = s.cop[k];
            if (j == 0) {
                if (op == 1) root_gt = true;
                else if (op == 0) return;  // candidates below the root: not this path
                continue;
            }
            if (op == 1) { L.gt_mask[i] |= 1u << j; lt.push_back({j, i}); }
            else if (op == 0) { L.lt_mask[i] |= 1u << j; lt.push_back({i, j}); }
            else L.ne_mask[i] |= 1u << j;
        }
        if (!root_bn || !root_gt) return;
        lt.push_back({0, i});
    }
    if (!complete_symmetry_breaking(n, adj, lt)) return;
    c->lprog = L;
    c->local_ok = true;
}

template <int N>
static void launch_local(uint64_t blocks, uint32_t bw, hipStream_t s, const LocalArgs& a) {
    hipLaunchKernelGGL(gm_local_kernel<N>, dim3(blocks), dim3(bw * kWave), 0, s, a);
}

int gm_local_run(g2_ctx* c, uint64_t* count, g2_stats* st) {
    const uint32_t bw = c->tune.block_waves;
    const uint64_t nblocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * waves_per_cu(c, WK_GM_LOCAL) / bw);
    const uint64_t nwaves = nblocks * bw;
    const uint64_t sw = c->tune.wave_scratch_words;
    int rc;
    if ((rc = ensure(c, c->counters, LCTR_N * 8, "counters"))) return rc;
    if ((rc = ensure(c, c->scratch, nwaves * sw * 8, "wave arenas"))) return rc;
    if ((rc = ensure(c, c->big_list2, std::max<uint64_t>(c->nv, 1) * 4, "deferred roots"))) return rc;
    hipError_t e;
    auto* ctr = (unsigned long long*)c->counters.ptr;
    if ((e = hipMemsetAsync(ctr, 0, LCTR_N * 8, c->stream))) return hip_err(c, e, "memset counters");
    LocalArgs a{};
    a.rp = (const uint64_t*)c->rp.ptr;
    a.cols = (const uint32_t*)c->cols.ptr;
    a.deg = (const uint32_t*)c->deg.ptr;
    a.list = nullptr;
    a.ntasks = c->nv;
    a.prog = c->lprog;
    a.scratch = (uint64_t*)c->scratch.ptr;
    a.scratch_words = sw;
    a.big_list = (uint32_t*)c->big_list2.ptr;
    a.ctr = ctr;
    auto launch = [&](uint64_t blocks) {
        switch (c->lprog.n) {
            case 2: launch_local<2>(blocks, bw, c->stream, a); break;
            case 3: launch_local<3>(blocks, bw, c->stream, a); break;
            case 4: launch_local<4>(blocks, bw, c->stream, a); break;
            case 5: launch_local<5>(blocks, bw, c->stream, a); break;
            case 6: launch_local<6>(blocks, bw, c->stream, a); break;
            case 7: launch_local<7>(blocks, bw, c->stream, a); break;
            default: launch_local<8>(blocks, bw, c->stream, a); break;
        }
        return hipGetLastError();
    };
    float k1 = 0, k2 = 0, tot = 0;
    if ((e = hipEventRecord(c->ev[0], c->stream))) return hip_err(c, e, "event");
    if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
    if ((e = launch(nblocks))) return hip_err(c, e, "local launch");
    if ((e = hipEventRecord(c->ev[2], c->stream))) return hip_err(c, e, "event");
    unsigned long long h[LCTR_N];
    if ((e = hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H counters");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "local sync");
    (void)hipEventElapsedTime(&k1, c->ev[1], c->ev[2]);
    uint32_t launches = 1;
    const uint64_t nbig = h[LCTR_NBIG];
    if (nbig) {
        const uint64_t need = h[LCTR_MAXNEED];
        uint64_t bwaves = std::min<uint64_t>(std::max<uint32_t>(c->tune.big_waves, 1), nbig);
        while (bwaves > 1 && c->budget && c->used - c->big_scratch.bytes + bwaves * need * 8 > c->budget) bwaves /= 2;
        bwaves = (bwaves + bw - 1) / bw * bw;
        if ((rc = ensure(c, c->big_scratch, bwaves * need * 8, "large-root arenas"))) return rc;
        a.list = (const uint32_t*)c->big_list2.ptr;
        a.ntasks = nbig;
        a.scratch = (uint64_t*)c->big_scratch.ptr;
        a.scratch_words = need;
        a.big_list = nullptr;
        if ((e = hipMemsetAsync(ctr + LCTR_NEXT, 0, 8, c->stream))) return hip_err(c, e, "memset next");
        if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
        if ((e = launch(bwaves / bw))) return hip_err(c, e, "local large-root launch");
        if ((e = hipEventRecord(c->ev[2], c->stream))) return hip_err(c, e, "event");
        launches = 2;
    }
    if ((e = hipEventRecord(c->ev[3], c->stream))) return hip_err(c, e, "event");
    if ((e = hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H counters");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "local sync");
    if (nbig) (void)hipEventElapsedTime(&k2, c->ev[1], c->ev[2]);
    (void)hipEventElapsedTime(&tot, c->ev[0], c->ev[3]);
    if (h[LCTR_ERR]) return set_err(c, G2_EBUDGET, "a root did not fit the large-root arena");
    if (count) *count = h[LCTR_COUNT];
    if (st) {
        st->kernel_ms = (double)k1 + (double)k2;
        st->launches = launches;
        st->total_ms = tot;
        st->nodes = h[LCTR_NODES];
        st->alg_bytes = 0;
        st->big_roots = nbig;
        st->max_scratch_words = h[LCTR_MAXNEED];
    }
    return G2_OK;
}

}  // namespace g2
