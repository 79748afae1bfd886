// csrc/g2_gm_local.hip -- oriented local-bitmap matching for root-dominated plans.
//
// Applies when the plan's root (position 0) is a back neighbour of, and
// ordered below (M(u_i) > M(u_0)), every other position -- every k-clique
// plan among the reference's presets -- and every neighbour of a source has
// its full row in the view (the whole graph, or a view bin materialised with
// hop >= 2, view/view_bin.h:61-159).  Then each match lies in
// U(v0) = { w in N(v0) : w above v0 } and the plan's back-neighbour /
// condition tables become bit operations over U.
//
// Orientation: "above" compares the rank pi(v) of the key (degree, id)
// instead of the id.  The plan's condition pairs are a complete symmetry
// breaking (checked for every ranking of the query vertices when the plan is
// set), so for any total order of the data vertices exactly one member of
// each automorphism class of an embedding satisfies them: the total count
// over all roots is the same number the reference's id order gives
// (app_gmatch/gmatch.h:559-588), while |U| is bounded by O(sqrt(E)) instead
// of by the largest degree.
//
// A view bin counts the matches whose lowest-ranked vertex is one of its
// sources.  The ranks of a view are taken over its own degrees; a source and
// its neighbours have their full rows, so their degrees and hence their
// relative order are the whole graph's, and vertices without a row rank
// below every vertex with one and never enter a U.  Summed over bins that
// partition the vertices, the counts give the whole-graph total (the
// reference's per-bin partial counts, which root each match at its
// lowest-id vertex, are split differently; their sum is the same).
//
// The first run on an uploaded view builds the oriented relabelled CSR on the
// device (later runs on the same view reuse it):
//   pi      = rank of (degree, id) over all vertices (radix sort),
//   N+(x)   = { pi(w) : w in N(v), pi(w) > x } for x = pi(v), sorted.
// U(x) is then the row N+(x) itself, already in rank order, and the rows of
// the local graph come from the short oriented rows N+(u) of its members
// instead of their full rows (a hub sits high in the order, so its N+ is
// short while its full row is the longest).
//
// Per root x:
//   1. U = N+(x), m = |U|, bit index = position in U;
//   2. rows[i] = N(U_i) & U as m-bit sets, from N+(U_i) & U[i+1..m) (both
//      bits set per found pair), by one flattened work list over the members
//      (stream N+(U_i) against U, or look U up in N+(U_i) when that row is
//      long);
//   3. the matches are walked depth first from the position-1 candidates:
//      cand(i) = AND of the rows of its back neighbours & order ranges &
//      NON_EQUAL holes, recomputed word by word (no stacks); the last
//      position is a popcount over the words its order range leaves.
// The roots go through passes, each taking the roots the previous one
// deferred: one wave per root with the rows in the wave's LDS (|U| <= 149);
// one workgroup of 256, 512 or 1024 threads per root with the rows in the
// workgroup's LDS (|U| <= 448, 700, 1039); one wave per root with the rows in
// a global arena sized to the largest remaining root.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "g2_device.hpp"
#include "g2_internal.hpp"

namespace g2 {

// counters: next task, matches, nodes, (deferred roots, their largest need)
// per pass, largest need that fit no arena
enum { LCTR_NEXT = 0, LCTR_COUNT, LCTR_NODES, LCTR_DEFER, LCTR_ERR = LCTR_DEFER + 8, LCTR_N };

constexpr int kLocalBlockWaves = 4;     // gm_local_kernel: __launch_bounds__(256)
constexpr uint32_t kLRowWords = 448;    // per-wave LDS rows (u64 words)
constexpr uint32_t kLU = 256;           // per-wave LDS copy of U (u32)
// workgroup passes: threads, LDS row words, |U| limit, value buckets, and
// workgroups per CU (the LDS each takes allows no more)
constexpr int kB1Threads = 256, kB1RowWords = 3136, kB1U = 448, kB1Buckets = 2048, kB1PerCu = 4;     // 25 KiB rows: |U| 448 at stride 7
constexpr int kB2Threads = 512, kB2RowWords = 7700, kB2U = 700, kB2Buckets = 2048, kB2PerCu = 2;      // 60 KiB rows: |U| 700 at stride 11
constexpr int kB3Threads = 1024, kB3RowWords = 17664, kB3U = 1040, kB3Buckets = 2047, kB3PerCu = 1;  // 138 KiB rows: |U| 1039 at stride 17
constexpr uint32_t kRev = 0x80000000u;  // work-list flag: look the later members up in the row
constexpr int kItems = 8;               // work items per thread in flight in the workgroup build
constexpr int kItemsW = 2;              // work items per lane in flight in the wave build
constexpr uint32_t kOwnScan = 8;        // a round or slice with at most this many owners after its first scans them
constexpr int kRowUnroll = 8;           // row entries per lane in flight while building the oriented rows
constexpr uint32_t kFetchRun = 8;       // roots per counter fetch in the first pass (a sweep of 4..64 favoured 8)
constexpr uint32_t kWB = 256;           // value buckets over U for a root in the wave's LDS (|U| <= 149: u8 indices)
constexpr uint32_t kWBMin = 32;         // smaller U: a search of all of U is as short

struct LocalArgs {
    const uint64_t* __restrict__ orp;   // oriented rows over ranks
    const uint32_t* __restrict__ ocols;
    const uint32_t* __restrict__ list;  // roots of the large-root pass (nullable: roots = [0, ntasks))
    uint64_t ntasks;
    LocalProgram prog;
    uint64_t* scratch;
    uint64_t scratch_words;
    uint32_t* big_list;                 // deferred roots (null in the last pass)
    uint32_t defer_ctr;                 // LCTR_DEFER + 2 * pass: count of big_list, need next to it
    unsigned long long* ctr;
    unsigned long long* tri;            // triangle features: count per rank
    uint8_t* flag;                      // triangle features: flag per oriented entry
};

// u64 words between consecutive local rows: ceil(m/64) rounded up to odd, so
// that lanes reading the rows of different ranks spread over the LDS banks
__host__ __device__ inline uint32_t row_stride(uint64_t m) { return (uint32_t)((m + 63) / 64) | 1u; }

// u64 words of a root's local rows
__host__ __device__ inline uint64_t local_need(uint64_t m) { return m * row_stride(m); }

// set bit b of a local row word (the 32-bit half holding it: a 32-bit
// atomic takes one LDS bank where the 64-bit one takes two)
__device__ inline void set_row_bit(uint64_t* rows, uint64_t word, uint32_t b) {
    atomicOr((unsigned int*)&rows[word] + ((b >> 5) & 1u), 1u << (b & 31u));
}

// bits of word w at positions in [lo, hi)
__device__ inline uint64_t range_word(uint32_t lo, uint32_t hi, uint32_t w) {
    const uint32_t base = w * 64;
    const uint64_t a = lo <= base ? ~0ull : (lo >= base + 64 ? 0 : ~0ull << (lo - base));
    const uint64_t b = hi >= base + 64 ? ~0ull : (hi <= base ? 0 : (1ull << (hi - base)) - 1);
    return a & b;
}

// word w of the candidates of position i given the matched ranks M[1..i-1]:
// the back neighbours' rows, within [lo, hi) (the order conditions' range
// from cand_range, which also keeps the bits below m), minus NON_EQUAL holes
__device__ inline uint64_t cand_word(const LocalProgram& P, uint32_t i, const uint32_t* M,
                                     const uint64_t* rows, uint32_t W, uint32_t lo, uint32_t hi, uint32_t w) {
    uint64_t x = range_word(lo, hi, w);
    uint32_t b = P.bn_mask[i];
    while (b) {
        const uint32_t j = __builtin_ctz(b);
        b &= b - 1;
        x &= rows[(uint64_t)M[j] * W + w];
    }
    uint32_t q = P.ne_mask[i];
    while (q) {
        const uint32_t j = __builtin_ctz(q);
        q &= q - 1;
        if ((M[j] >> 6) == w) x &= ~(1ull << (M[j] & 63));
    }
    return x;
}

// candidate positions of position i lie in [*lo, *hi) by its order conditions
__device__ inline void cand_range(const LocalProgram& P, uint32_t i, const uint32_t* M, uint32_t m,
                                  uint32_t* lo, uint32_t* hi) {
    uint32_t a = 0, b = m;
    uint32_t g = P.gt_mask[i];
    while (g) {
        const uint32_t j = __builtin_ctz(g);
        g &= g - 1;
        a = max(a, M[j] + 1);
    }
    uint32_t l = P.lt_mask[i];
    while (l) {
        const uint32_t j = __builtin_ctz(l);
        l &= l - 1;
        b = min(b, M[j]);
    }
    *lo = a;
    *hi = b;
}

// Local rows of members a0 .. a0+63: pairs (i, j), i < j, with U_j in
// N+(U_i), both bits set.  One flattened work list keeps all lanes busy:
// an item is one entry of N+(U_i) (looked up in U), or, when that row is
// longer than (later members) x (its search depth), one later member U_j
// (looked up in N+(U_i), global memory).
// Each lane keeps kItemsW items in flight; U is searched by a fixed number
// of halving steps so the items' loads overlap.  rows and U are LDS or
// global memory (fixed per call site after inlining).  MODE 0 sets both
// bits of a found pair; MODE 1 (clique plans) only the later member's, so
// row i holds the members above i; MODE 2 sets no rows but counts the
// triangle (root, U_i, U_j): one count for each member in a.tri, one for the
// root in hits, and the flags of its three oriented entries (root row from
// entry xlo on).
// Value buckets over U (optional): bucket(v) = floor((v - ulo) * NB / range)
// by one high multiply (the identity when U spans at most NB ranks); a
// lookup of w searches U[bkt[b], bkt[b+1]) for b = bucket(w), in hb halving
// steps (hb: the highest power of two <= the largest bucket).
struct UBuckets {
    const uint8_t* bkt;  // null: search all of U
    uint32_t nb, ulo, bmul, hb;
    uint32_t imask;      // ~0 when U spans at most nb ranks (then bmul = 0 and the bucket is v - ulo), else 0
    __device__ uint32_t bucket(uint32_t v) const {
        const uint32_t d = v - ulo;
        return min(__umulhi(d, bmul) + (d & imask), nb - 1);
    }
};

template <int MODE>
__device__ __forceinline__ void build_chunk(const LocalArgs& a, const uint32_t* U, uint64_t* rows, uint32_t m,
                                            uint32_t W, uint32_t a0, uint64_t xlo, uint32_t& hits,
                                            const UBuckets& B) {
    const uint32_t lane = lane_id();
    const uint32_t na = umin32(kWave, m - 1 - a0);  // members with a later member
    uint32_t st = 0;  // (row offsets fit 32 bits: this path runs with fewer than 2^32 row entries)
    uint32_t code = 0, len = 0;
    if (lane < na) {
        const uint32_t ia = a0 + lane, u = U[ia];
        uint64_t e0 = a.orp[u], e1 = a.orp[u + 1];
        const uint64_t rem = m - 1 - ia;
        // a long row whose later members are few is searched whole (no
        // restricting searches on the chain); otherwise it is restricted to
        // [U[ia+1], U[m-1]] and the cheaper side chosen again
        bool rev = rem * (1 + bitlen(e1 - e0)) < e1 - e0;
        if (!rev && e1 - e0 > kWave) {
            lower_bound2(a.ocols, e0, e1, U[ia + 1], U[m - 1] + 1, &e0, &e1);
            rev = rem * (1 + bitlen(e1 - e0)) < e1 - e0;
        }
        const uint64_t n = e1 - e0;
        rev = rev || n > (1u << 20);
        st = (uint32_t)e0;
        code = rev ? ((uint32_t)n | kRev) : (uint32_t)n;
        len = rev ? (uint32_t)rem : (uint32_t)n;
    }
    uint32_t incl = len;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, kWave);
        if (lane >= (uint32_t)o) incl += y;
    }
    const uint32_t pref = incl - len;
    const uint32_t T = read_lane(incl, kWave - 1);
    uint32_t hb_u = 1;  // highest power of two <= m
    while (hb_u * 2 <= m) hb_u *= 2;
    for (uint32_t f0 = 0; f0 < T; f0 += kItemsW * kWave) {
        uint32_t f[kItemsW], j[kItemsW], item[kItemsW], cj[kItemsW], w[kItemsW], pos[kItemsW];
        uint32_t sj[kItemsW];
#pragma unroll
        for (int q = 0; q < kItemsW; q++) {
            f[q] = f0 + q * kWave + lane;
            j[q] = 0;
        }
        // owners: the last member whose items start at or before f.  The
        // round's first and last owners come from two ballots (pref is
        // non-decreasing over the lanes); each item then searches only between
        // them (every lane runs the same steps, so the lane reads see all
        // lanes active), and a round inside one member needs no search
        const uint32_t flast = umin32(f0 + kItemsW * kWave, T) - 1;
        const uint32_t o0 = __popcll(ballot(lane < na && pref <= f0)) - 1;
        const uint32_t o1 = __popcll(ballot(lane < na && pref <= flast)) - 1;
#pragma unroll
        for (int q = 0; q < kItemsW; q++) j[q] = o0;
        if (o1 > o0 && o1 - o0 <= kOwnScan) {
            // few owners: count the member starts at or below each item, the
            // starts read once per round as scalars
            for (uint32_t o = o0 + 1; o <= o1; o++) {
                const uint32_t po = read_lane(pref, o);
#pragma unroll
                for (int q = 0; q < kItemsW; q++) j[q] += po <= f[q] ? 1u : 0u;
            }
        } else if (o1 > o0) {
            uint32_t hb_o = 1;
            while (hb_o * 2 <= o1 - o0) hb_o *= 2;
            for (uint32_t step = hb_o; step >= 1; step >>= 1) {
#pragma unroll
                for (int q = 0; q < kItemsW; q++) {
                    const uint32_t c = j[q] + step;
                    const uint32_t pc = __shfl(pref, (int)umin32(c, o1), kWave);
                    if (c <= o1 && pc <= f[q]) j[q] = c;
                }
            }
        }
        uint32_t hb_r = 0;  // largest search range among this lane's lookups
#pragma unroll
        for (int q = 0; q < kItemsW; q++) {
            if (o1 == o0) {  // (wave-uniform: one member owns the whole round)
                item[q] = f[q] - read_lane(pref, o0);
                cj[q] = read_lane(code, o0);
                sj[q] = read_lane(st, o0);
            } else {
                item[q] = f[q] - __shfl(pref, (int)j[q], kWave);
                cj[q] = __shfl(code, (int)j[q], kWave);
                sj[q] = __shfl(st, (int)j[q], kWave);
            }
            const bool valid = f[q] < T, rev = valid && (cj[q] & kRev);
            hb_r = max(hb_r, rev ? cj[q] & ~kRev : 0u);
            // both candidates are loaded without branching (no lane waits on a
            // load another path issued): the later member to look up, and the
            // row entry (a reversed or invalid lane reads entry xlo, inside the
            // array, instead of a row entry it does not use)
            pos[q] = U[umin32(a0 + j[q] + 1 + item[q], m - 1)];
            w[q] = a.ocols[valid && !rev ? (uint64_t)sj[q] + item[q] : xlo];
        }
#pragma unroll
        for (int q = 0; q < kItemsW; q++) {
            w[q] = f[q] >= T ? 0u : (cj[q] & kRev) ? pos[q] : w[q];
            pos[q] = 0;
        }
        uint32_t lim[kItemsW];
        uint32_t hb = hb_u;
#pragma unroll
        for (int q = 0; q < kItemsW; q++) lim[q] = m;
        if (B.bkt) {  // (an entry above U_{m-1} lands in the last bucket and misses)
            hb = B.hb;
#pragma unroll
            for (int q = 0; q < kItemsW; q++) {
                const uint32_t b = B.bucket(w[q]);
                pos[q] = B.bkt[b];
                lim[q] = B.bkt[b + 1];
            }
        }
        for (uint32_t step = hb; step >= 1; step >>= 1) {  // stream: pos = |{U_j < w}|
#pragma unroll
            for (int q = 0; q < kItemsW; q++) {  // (clamped read, no branch)
                const uint32_t c = pos[q] + step;
                const uint32_t v = U[min(c, m) - 1];
                pos[q] = (!(cj[q] & kRev) && c <= lim[q] && v < w[q]) ? c : pos[q];
            }
        }
        // lookups: rpos = |{entries <= w}| in the member's row, with the entry
        // at rpos-1 kept from the step that accepted it
        uint32_t rpos[kItemsW], last[kItemsW];
#pragma unroll
        for (int q = 0; q < kItemsW; q++) {
            rpos[q] = 0;
            last[q] = ~0u;
        }
        uint32_t step0 = 1;
        while (step0 * 2 <= hb_r) step0 *= 2;
        for (uint32_t step = hb_r ? step0 : 0; step >= 1; step >>= 1) {
#pragma unroll
            for (int q = 0; q < kItemsW; q++) {
                const uint32_t c = rpos[q] + step;
                if ((cj[q] & kRev) && c <= (cj[q] & ~kRev)) {
                    const uint32_t v = a.ocols[(uint64_t)sj[q] + c - 1];
                    if (v <= w[q]) {
                        rpos[q] = c;
                        last[q] = v;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kItemsW; q++) {
            if (f[q] >= T) continue;
            const uint32_t ia = a0 + j[q];
            uint32_t ib;
            bool hit;
            if (cj[q] & kRev) {
                ib = ia + 1 + item[q];
                hit = rpos[q] > 0 && last[q] == w[q];
            } else {
                ib = pos[q];
                const uint32_t uv = U[min(ib, m - 1)];
                hit = ib < m && uv == w[q];
            }
            if (hit) {
                if constexpr (MODE == 2) {
                    hits++;
                    atomicAdd(&a.tri[U[ia]], 1ull);
                    atomicAdd(&a.tri[U[ib]], 1ull);
                    a.flag[xlo + ia] = 1;
                    a.flag[xlo + ib] = 1;
                    a.flag[(cj[q] & kRev) ? (uint64_t)sj[q] + rpos[q] - 1 : (uint64_t)sj[q] + item[q]] = 1;
                } else {
                    set_row_bit(rows, (uint64_t)ia * W + (ib >> 6), ib);
                    if constexpr (MODE == 0)
                        set_row_bit(rows, (uint64_t)ib * W + (ia >> 6), ia);
                }
            }
        }
    }
}

// Clique plans (position i adjacent to every earlier one and above i-1):
// with rows that hold only later members, the candidates of position L are
// the AND of the matched members' rows, from the word of the last one on --
// no order range, no holes.  off[1..L-1] are the matched rows' word offsets,
// last the rank at position L-1, nw the words that hold bits.
template <int N, int L>
__device__ __forceinline__ void clique_walk(const uint64_t* rows, uint32_t W, uint32_t nw, uint32_t* off, uint32_t last,
                                            uint64_t& lane_total, uint64_t& nodes) {
    if constexpr (L >= N - 1) {
        uint32_t c = 0;
        for (uint32_t w = last >> 6; w < nw; w++) {
            uint64_t x = rows[off[1] + w];
#pragma unroll
            for (int j = 2; j < L; j++) x &= rows[off[j] + w];
            c += (uint32_t)__popcll(x);
        }
        lane_total += c;
    } else {
        for (uint32_t w = last >> 6; w < nw; w++) {
            uint64_t x = rows[off[1] + w];
#pragma unroll
            for (int j = 2; j < L; j++) x &= rows[off[j] + w];
            while (x) {
                const uint32_t b = w * 64 + (uint32_t)__builtin_ctzll(x);
                x &= x - 1;
                nodes++;
                off[L] = b * W;
                clique_walk<N, L + 1>(rows, W, nw, off, b, lane_total, nodes);
            }
        }
    }
}

// This lane's depth-first walk of positions level0 .. N-1 given the matched
// ranks M[1 .. level0-1]: cand(i) is recomputed word by word from the rows
// (no stacks), the last position is a popcount over its order range.
template <int N>
__device__ __forceinline__ void walk_from(const LocalProgram& P, uint32_t* M, uint32_t level0, const uint64_t* rows,
                                          uint32_t W, uint32_t m, uint64_t& lane_total, uint64_t& nodes) {
    uint32_t cur[N], hiv[N];
    uint32_t plo, phi;
    cand_range(P, level0, M, m, &plo, &phi);
    if (level0 == N - 1) {
        if (plo < phi)
            for (uint32_t w = plo >> 6; w <= (phi - 1) >> 6; w++) lane_total += __popcll(cand_word(P, level0, M, rows, W, plo, phi, w));
        return;
    }
    uint32_t level = level0;
    cur[level] = plo;
    hiv[level] = phi;
    while (true) {
        // next member of cand(level) at or after cur[level]
        uint32_t pos = cur[level];
        const uint32_t hi = hiv[level];
        bool found = false;
        while (pos < hi) {
            const uint32_t w = pos >> 6;
            const uint64_t x = cand_word(P, level, M, rows, W, 0, hi, w) & (~0ull << (pos & 63));  // (pos >= the range's low end)
            if (x) {
                pos = w * 64 + __builtin_ctzll(x);
                found = pos < hi;
                break;
            }
            pos = (w + 1) * 64;
        }
        if (!found) {
            if (level == level0) break;
            level--;
            continue;
        }
        cur[level] = pos + 1;
        M[level] = pos;
        const uint32_t nx = level + 1;
        nodes++;
        cand_range(P, nx, M, m, &plo, &phi);
        if (nx == N - 1) {
            if (plo < phi)
                for (uint32_t w = plo >> 6; w <= (phi - 1) >> 6; w++) lane_total += __popcll(cand_word(P, nx, M, rows, W, plo, phi, w));
        } else {
            cur[nx] = plo;
            hiv[nx] = phi;
            level = nx;
        }
    }
}

// One root on one wave: build the local rows over U, then each lane walks the
// matches of the position-1 candidates lane, lane+64, ...
template <int N, bool CLQ>
__device__ __forceinline__ void local_root(const LocalArgs& a, uint64_t lo, uint32_t m, uint64_t* rows,
                                           const uint32_t* U, uint32_t* Ucopy, uint8_t* bkt, uint64_t& lane_total,
                                           uint64_t& nodes) {
    const uint32_t lane = lane_id();
    const uint32_t W = row_stride(m);
    for (uint32_t i = lane; i < m * W; i += kWave) rows[i] = 0;
    if (Ucopy)
        for (uint32_t i = lane; i < m; i += kWave) Ucopy[i] = a.ocols[lo + i];
    wave_sync();
    UBuckets B{};
    if (bkt && m > kWBMin) {  // value buckets over U (U in LDS, m <= 149)
        const uint32_t range = U[m - 1] - U[0] + 1;
        B.nb = kWB;
        B.ulo = U[0];
        // (uniform values made opaque to the compiler: it would otherwise
        // widen the high multiply to 64 bits)
        B.imask = __builtin_amdgcn_readfirstlane(range <= kWB ? ~0u : 0u);
        B.bmul = __builtin_amdgcn_readfirstlane(range <= kWB ? 0u : (uint32_t)(((uint64_t)kWB << 32) / range));
        for (uint32_t i = lane; i <= m; i += kWave) {  // buckets (bucket(U_{i-1}), bucket(U_i)] start at i
            const uint32_t bi = i < m ? B.bucket(U[i]) : kWB;
            for (uint32_t b = i == 0 ? 0 : B.bucket(U[i - 1]) + 1; b <= bi; b++) bkt[b] = (uint8_t)i;
        }
        wave_sync();
        uint32_t big = 0;
        for (uint32_t b = lane; b < kWB; b += kWave) big = umax32(big, bkt[b + 1] - bkt[b]);
        big = (uint32_t)wave_max((uint64_t)big);
        B.hb = 1;
        while (B.hb * 2 <= big) B.hb *= 2;
        B.bkt = bkt;
    }
    for (uint32_t a0 = 0; a0 + 1 < m; a0 += kWave) {
        uint32_t unused = 0;
        build_chunk<CLQ ? 1 : 0>(a, U, rows, m, W, a0, 0, unused, B);
        wave_sync();  // (the chunk's atomics complete before the walk reads rows)
    }
    if constexpr (CLQ) {
        const uint32_t nw = (m + 63) >> 6;
        uint32_t off[N];
        for (uint32_t r1 = lane; r1 < m; r1 += kWave) {
            nodes++;
            off[1] = r1 * W;
            clique_walk<N, 2>(rows, W, nw, off, r1, lane_total, nodes);
        }
    } else {
        uint32_t M[N];
        for (uint32_t r1 = lane; r1 < m; r1 += kWave) {
            M[1] = r1;
            nodes++;
            walk_from<N>(a.prog, M, 2, rows, W, m, lane_total, nodes);
        }
    }
}

// put root x on the deferred list of the next pass
__device__ inline void defer_root(const LocalArgs& a, uint32_t x, uint64_t need) {
    if (a.big_list) {
        const unsigned long long idx = atomicAdd(&a.ctr[a.defer_ctr], 1ull);
        a.big_list[idx] = x;
        atomicMax(&a.ctr[a.defer_ctr + 1], (unsigned long long)need);
    } else {
        atomicMax(&a.ctr[LCTR_ERR], (unsigned long long)need);
    }
}

// Pass 1 (big_list set): roots whose rows fit the wave's LDS; the others are
// deferred.  Pass 3 (big_list null): the roots too large for a block's LDS,
// rows in per-wave global arenas of scratch_words.
template <int N, bool CLQ>
__global__ __launch_bounds__(256, (N <= 4 ? 8 : 4)) void gm_local_kernel(LocalArgs a) {
    __shared__ uint64_t lrows[kLocalBlockWaves][kLRowWords];
    __shared__ uint32_t lU[kLocalBlockWaves][kLU];
    __shared__ uint8_t lB[kLocalBlockWaves][kWB + 4];
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / kWave;
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + wid;
    uint64_t* ws = a.scratch + gwave * a.scratch_words;
    uint64_t lane_total = 0, nodes = 0;

    // roots come in runs of kFetchRun from the counter in the first pass (most
    // of them are trivial, and one atomic per root would queue the waves on
    // the counter), singly in the third (few, heavy roots).  The first pass
    // takes its roots from the top rank down: the roots with the largest
    // local graphs sit in the upper middle of the (degree, id) order and the
    // low ranks are the cheapest, so the pass ends on short roots instead of
    // on a tail of long ones (14.5 -> 13.7 ms on R-MAT 22, 4-cliques).
    const uint32_t run = a.big_list ? kFetchRun : 1;
    uint64_t t = 0, tend = 0;
    while (true) {
        if (t == tend) {
            t = fetch_tasks(&a.ctr[LCTR_NEXT], run);
            tend = umin64(t + run, a.ntasks);
        }
        if (t >= a.ntasks) break;
        const uint64_t tt = a.big_list ? a.ntasks - 1 - t : t;
        const uint32_t x = uniform(a.list ? a.list[tt] : (uint32_t)tt);
        t++;
        const uint64_t lo = uniform(a.orp[x]);
        const uint32_t m = (uint32_t)(uniform(a.orp[x + 1]) - lo);
        if (m < (uint32_t)(N - 1)) continue;
        if constexpr (N == 2) {
            if (lane == 0) lane_total += m;
            continue;
        } else {
            const uint64_t need = local_need(m);
            if (m <= kLU && need <= kLRowWords) {
                local_root<N, CLQ>(a, lo, m, lrows[wid], lU[wid], lU[wid], lB[wid], lane_total, nodes);
            } else if (!a.big_list && need <= a.scratch_words) {
                local_root<N, CLQ>(a, lo, m, ws, a.ocols + lo, nullptr, nullptr, lane_total, nodes);
            } else if (lane == 0) {
                defer_root(a, x, need);
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

// Pass 2: one root per workgroup of BT threads, its rows in the workgroup's
// LDS (up to ROWW words, |U| <= BU); roots beyond that are deferred to the
// next pass.  Three shapes run in turn: 256 threads with 25 KB of rows (four
// workgroups per CU, |U| <= 448), 512 threads with 60 KB (two per CU, |U| <=
// 700), then 1024 threads with 138 KB (one per CU).
// The build is one work list over all members and threads; then waves take
// position-1 candidates r1 one at a time from an LDS counter, a wave's lanes
// share the position-2 candidates of its r1, queued 64 at a time in LDS, and
// each lane walks the rest of its (r1, r2) subtree.
template <int N, bool CLQ, int BT, int ROWW, int BU, int NB>
__global__ __launch_bounds__(BT, 4) void gm_local_block_kernel(LocalArgs a) {
    constexpr int kWaves = BT / kWave;
    constexpr int kMembersPerThread = (BU + BT - 1) / BT;
    __shared__ uint64_t rows[ROWW];
    __shared__ uint32_t U[BU];
    __shared__ uint32_t mpref[BU];   // members' first work item
    __shared__ uint32_t mstart[BU];  // members' first entry in ocols
    __shared__ uint32_t mcode[BU];   // members' entry count | kRev
    __shared__ uint32_t wsum[kWaves];
    __shared__ unsigned long long s_task;
    __shared__ uint32_t s_next, s_maxb;
    // the build's value-bucket index (first index of U in each bucket) and the
    // count's per-wave candidate queues share one buffer: a barrier separates
    // the build from the count, and the next root's index from this count
    constexpr int kQueueBytes = kWaves * kWave * 4, kBucketBytes = (NB + 1) * 2;
    __shared__ __align__(16) uint8_t qb_raw[kQueueBytes > kBucketBytes ? kQueueBytes : kBucketBytes];
    auto queue = reinterpret_cast<uint32_t (*)[kWave]>(qb_raw);
    uint16_t* const ubkt = reinterpret_cast<uint16_t*>(qb_raw);
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t nwaves = blockDim.x / kWave;
    const uint32_t tid = threadIdx.x;
    const LocalProgram& P = a.prog;
    uint64_t lane_total = 0, nodes = 0;

    while (true) {
        __syncthreads();  // every wave is done with the previous root's rows
        if (threadIdx.x == 0) {
            s_task = atomicAdd(&a.ctr[LCTR_NEXT], 1ull);
            s_next = 0;
        }
        __syncthreads();
        const uint64_t t = s_task;
        if (t >= a.ntasks) break;
        const uint32_t x = a.list[t];
        const uint64_t lo = a.orp[x];
        const uint32_t m = (uint32_t)(a.orp[x + 1] - lo);
        const uint32_t W = row_stride(m);
        const uint64_t need = local_need(m);
        if (m < (uint32_t)(N - 1)) continue;
        if (m > BU || need > ROWW) {
            if (threadIdx.x == 0) defer_root(a, x, need);
            continue;
        }
        for (uint32_t i = tid; i < need; i += blockDim.x) rows[i] = 0;
        for (uint32_t i = tid; i < m; i += blockDim.x) U[i] = a.ocols[lo + i];
        if (tid == 0) s_maxb = 0;
        __syncthreads();
        // value buckets over [U_0, U_{m-1}]: a lookup searches one bucket
        const uint32_t ulo = U[0];
        // bucket(v) = floor((v - ulo) * NB / range) by one 32-bit high multiply
        // (the identity when U spans at most NB ranks); monotone, and below NB
        // on [ulo, U_{m-1}]
        const uint32_t range = U[m - 1] - ulo + 1;
        // (uniform values made opaque to the compiler: it would otherwise
        // widen the high multiply to 64 bits)
        const uint32_t imask = __builtin_amdgcn_readfirstlane(range <= (uint32_t)NB ? ~0u : 0u);
        const uint32_t bmul =
            __builtin_amdgcn_readfirstlane(range <= (uint32_t)NB ? 0u : (uint32_t)(((uint64_t)NB << 32) / range));
        auto bucket = [&](uint32_t v) {
            const uint32_t d = v - ulo;
            return __umulhi(d, bmul) + (d & imask);
        };
        for (uint32_t i = tid; i <= m; i += blockDim.x) {  // buckets (bucket(U_{i-1}), bucket(U_i)] start at i
            const uint32_t bi = i < m ? bucket(U[i]) : NB;
            for (uint32_t b = i == 0 ? 0 : bucket(U[i - 1]) + 1; b <= bi; b++) ubkt[b] = (uint16_t)i;
        }
        __syncthreads();
        {
            uint32_t big = 0;
            for (uint32_t b = tid; b < NB; b += blockDim.x) big = umax32(big, ubkt[b + 1] - ubkt[b]);
            big = (uint32_t)wave_max((uint64_t)big);
            if (lane == 0 && big) atomicMax(&s_maxb, big);
        }
        // ---- local rows: one work list over all members and threads ---------
        // member i < m-1 has either n entries of N+(U_i) (restricted to
        // [U_{i+1}, U_{m-1}]) to look up in U (LDS), or, when n exceeds
        // (m-1-i) x (1 + log2 n), its m-1-i later members to look up in
        // N+(U_i) (global).
        // Thread t describes members t, t + BT, ...; their item counts go to
        // mpref, which a block scan then turns into first items.
#pragma unroll
        for (int h = 0; h < kMembersPerThread; h++) {
            const uint32_t i = tid + h * BT;
            if (i + 1 < m) {
                const uint32_t u = U[i];
                uint64_t e0 = a.orp[u], e1 = a.orp[u + 1];
                if (e1 - e0 > kWave) lower_bound2(a.ocols, e0, e1, U[i + 1], U[m - 1] + 1, &e0, &e1);
                const uint64_t n = e1 - e0, rem = m - 1 - i;
                const bool rev = n > (1u << 20) || rem * (1 + bitlen(n)) < n;
                mstart[i] = (uint32_t)e0;
                mcode[i] = rev ? ((uint32_t)n | kRev) : (uint32_t)n;
                mpref[i] = rev ? (uint32_t)rem : (uint32_t)n;
            } else if (i < m) {
                mpref[i] = 0;
            }
        }
        __syncthreads();
        uint32_t T = 0;
        {  // exclusive scan of mpref[0..m) in member order, contiguous runs per thread
            const uint32_t per = (m + BT - 1) / BT;
            const uint32_t c0 = min(tid * per, m), c1 = min(c0 + per, m);
            uint32_t sum = 0;
            for (uint32_t i = c0; i < c1; i++) sum += mpref[i];
            uint32_t incl = sum;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o, kWave);
                if (lane >= (uint32_t)o) incl += y;
            }
            if (lane == kWave - 1) wsum[wid] = incl;
            __syncthreads();
            uint32_t before = 0;
            for (uint32_t q = 0; q < nwaves; q++) {
                const uint32_t ws = wsum[q];
                before += q < wid ? ws : 0;
                T += ws;
            }
            uint32_t run = before + incl - sum;
            for (uint32_t i = c0; i < c1; i++) {
                const uint32_t l = mpref[i];
                mpref[i] = run;
                run += l;
            }
        }
        __syncthreads();
        const uint32_t nm = m - 1;  // members that own items
        uint32_t hb_own = 1, hb_u = 1;  // highest powers of two <= nm-1 and <= the largest bucket
        while (hb_own * 2 <= nm - 1) hb_own *= 2;
        while (hb_u * 2 <= s_maxb) hb_u *= 2;
        constexpr uint32_t kSlice = kItems * kWave;  // consecutive items per wave and round
        for (uint32_t f0 = wid * kSlice; f0 < T; f0 += nwaves * kSlice) {
            // the slice's first and last owners (wave-uniform searches, broadcast
            // reads), then each lane searches only between them
            const uint32_t flast = min(f0 + kSlice, T) - 1;
            uint32_t o0 = 0, o1 = 0;
            for (uint32_t step = hb_own; step >= 1; step >>= 1) {
                if (o0 + step < nm && mpref[o0 + step] <= f0) o0 += step;
                if (o1 + step < nm && mpref[o1 + step] <= flast) o1 += step;
            }
            o0 = __builtin_amdgcn_readfirstlane(o0);  // (wave-uniform: scalar from here)
            o1 = __builtin_amdgcn_readfirstlane(o1);
            uint32_t hb_r = 1;
            while (hb_r * 2 <= o1 - o0) hb_r *= 2;
            uint32_t f[kItems], own[kItems], w[kItems], pos[kItems];
#pragma unroll
            for (int q = 0; q < kItems; q++) {
                f[q] = f0 + q * kWave + lane;
                own[q] = o0;
            }
            if (o1 > o0 && o1 - o0 <= kOwnScan) {
                // few owners: count the member starts at or below each item,
                // each start read once per slice as a scalar
                for (uint32_t o = o0 + 1; o <= o1; o++) {
                    const uint32_t po = __builtin_amdgcn_readfirstlane(mpref[o]);
#pragma unroll
                    for (int q = 0; q < kItems; q++) own[q] += po <= f[q] ? 1u : 0u;
                }
            } else if (o1 > o0) {
                for (uint32_t step = hb_r; step >= 1; step >>= 1) {
#pragma unroll
                    for (int q = 0; q < kItems; q++) {
                        const uint32_t c = own[q] + step;
                        if (c <= o1 && mpref[c] <= f[q]) own[q] = c;
                    }
                }
            }
            // w: the entry to look up in U, or for a reversed member the later
            // member to look up in the member's row (rlen > 0: that row's length)
            uint32_t rst[kItems], rlen[kItems], rmax = 0;
            // every lane loads both candidates of all its items before any is
            // used (no branch around a load, so no lane waits on a load that
            // another path issued): a reversed member's row is longer than its
            // later-member count, so rst + item stays inside it, and an invalid
            // lane reads the root's own row
#pragma unroll
            for (int q = 0; q < kItems; q++) {
                const uint32_t o = own[q], code = mcode[o], item = f[q] - mpref[o];
                const bool valid = f[q] < T;
                rst[q] = mstart[o];
                rlen[q] = valid && (code & kRev) ? (code & ~kRev) : 0;
                rmax = max(rmax, rlen[q]);
                pos[q] = U[umin32(o + 1 + item, m - 1)];
                w[q] = a.ocols[valid ? (uint64_t)rst[q] + item : lo];
            }
#pragma unroll
            for (int q = 0; q < kItems; q++) w[q] = f[q] >= T ? ulo : rlen[q] ? pos[q] : w[q];
            uint32_t lim[kItems];
#pragma unroll
            for (int q = 0; q < kItems; q++) {  // (an entry above U_{m-1} lands in the last bucket and misses)
                const uint32_t b = umin32(bucket(w[q]), NB - 1);
                pos[q] = ubkt[b];
                lim[q] = ubkt[b + 1];
            }
            for (uint32_t step = hb_u; step >= 1; step >>= 1) {  // pos = |{U_j < w}|, within w's bucket
#pragma unroll
                for (int q = 0; q < kItems; q++) {
                    const uint32_t c = pos[q] + step;
                    if (c <= lim[q] && U[c - 1] < w[q]) pos[q] = c;
                }
            }
            // reversed members: rpos = |{entries <= w}| in the member's row, the
            // entry at rpos-1 kept; the items' searches interleave step by step
            // so their global loads overlap
            uint32_t rpos[kItems], last[kItems];
#pragma unroll
            for (int q = 0; q < kItems; q++) {
                rpos[q] = 0;
                last[q] = ~0u;
            }
            uint32_t rstep = 1;
            while (rstep * 2 <= rmax) rstep *= 2;
            for (uint32_t step = rmax ? rstep : 0; step >= 1; step >>= 1) {
#pragma unroll
                for (int q = 0; q < kItems; q++) {
                    const uint32_t c = rpos[q] + step;
                    if (c <= rlen[q]) {
                        const uint32_t v = a.ocols[(uint64_t)rst[q] + c - 1];
                        if (v <= w[q]) {
                            rpos[q] = c;
                            last[q] = v;
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < kItems; q++) {
                if (f[q] >= T) continue;
                const uint32_t ia = own[q];
                uint32_t ib;
                bool hit;
                if (rlen[q]) {
                    ib = ia + 1 + (f[q] - mpref[ia]);
                    hit = last[q] == w[q];
                } else {
                    ib = pos[q];
                    hit = ib < m && U[ib] == w[q];
                }
                if (hit) {  // (row offsets within the LDS rows fit 32 bits)
                    set_row_bit(rows, ia * W + (ib >> 6), ib);
                    if constexpr (!CLQ)
                        set_row_bit(rows, ib * W + (ia >> 6), ia);
                }
            }
        }
        __syncthreads();

        if constexpr (CLQ) {
            // position-1 members r1 one at a time per wave; the (r1, r2)
            // pairs of successive r1 share the wave's queue, 64 at a time
            const uint32_t nw = (m + 63) >> 6;
            uint32_t k = 0;
            auto drain_pairs = [&](uint32_t nq) {
                wave_sync();
                if (lane < nq) {
                    const uint32_t e = queue[wid][lane];
                    uint32_t off[N];
                    off[1] = (e >> 16) * W;
                    off[2] = (e & 0xffffu) * W;
                    nodes++;
                    clique_walk<N, 3>(rows, W, nw, off, e & 0xffffu, lane_total, nodes);
                }
                wave_sync();
            };
            while (true) {
                __builtin_amdgcn_wave_barrier();
                uint32_t r1 = 0;
                if (lane == 0) r1 = atomicAdd(&s_next, 1u);
                r1 = uniform(r1);
                if (r1 >= m) break;
                if (lane == 0) nodes++;
                if constexpr (N == 3) {
                    for (uint32_t w = (r1 >> 6) + lane; w < nw; w += kWave) lane_total += __popcll(rows[r1 * W + w]);
                } else {
                    for (uint32_t w = r1 >> 6; w < nw; w++) {
                        const uint64_t cw = rows[r1 * W + w];
                        const uint32_t cnt = (uint32_t)__popcll(cw);
                        if (k + cnt > kWave) {
                            drain_pairs(k);
                            k = 0;
                        }
                        if ((cw >> lane) & 1) queue[wid][k + lanes_below(cw)] = (r1 << 16) | (w * 64 + lane);
                        k += cnt;
                    }
                }
            }
            if constexpr (N > 3) {
                if (k) drain_pairs(k);
            }
            continue;
        }
        uint32_t M[N];
        while (true) {
            __builtin_amdgcn_wave_barrier();
            uint32_t r1 = 0;
            if (lane == 0) r1 = atomicAdd(&s_next, 1u);
            r1 = uniform(r1);
            if (r1 >= m) break;
            M[1] = r1;
            if (lane == 0) nodes++;
            uint32_t plo, phi;
            cand_range(P, 2, M, m, &plo, &phi);
            if (plo >= phi) continue;
            if constexpr (N == 3) {
                for (uint32_t w = (plo >> 6) + lane; w <= (phi - 1) >> 6; w += kWave)
                    lane_total += __popcll(cand_word(P, 2, M, rows, W, plo, phi, w));
            } else {
                // position-2 candidates of r1, 64 at a time
                uint32_t k = 0;
                auto drain_queue = [&](uint32_t nq) {
                    wave_sync();
                    if (lane < nq) {
                        M[2] = queue[wid][lane];
                        nodes++;
                        walk_from<N>(P, M, 3, rows, W, m, lane_total, nodes);
                    }
                    wave_sync();
                };
                for (uint32_t w = plo >> 6; w <= (phi - 1) >> 6; w++) {
                    const uint64_t cw = cand_word(P, 2, M, rows, W, plo, phi, w);
                    const uint32_t cnt = (uint32_t)__popcll(cw);
                    if (k + cnt > kWave) {
                        drain_queue(k);
                        k = 0;
                    }
                    if ((cw >> lane) & 1) queue[wid][k + lanes_below(cw)] = w * 64 + lane;
                    k += cnt;
                }
                if (k) drain_queue(k);
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

// ---- oriented relabelled CSR ------------------------------------------------

__global__ __launch_bounds__(256) void okey_kernel(const uint32_t* __restrict__ deg, uint64_t nv,
                                                   unsigned long long* __restrict__ keys) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (uint64_t)gridDim.x * blockDim.x)
        keys[i] = ((unsigned long long)deg[i] << 32) | i;
}

__global__ __launch_bounds__(256) void opi_kernel(const unsigned long long* __restrict__ sorted, uint64_t nv,
                                                  uint32_t* __restrict__ pi) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (uint64_t)gridDim.x * blockDim.x)
        pi[(uint32_t)sorted[i]] = (uint32_t)i;
}

// Out-degree of every rank, with the above-neighbours' ranks of vertex v
// compacted into tmp[rp[v] ..] (row order): one wave per vertex, kRowUnroll
// row entries per lane in flight.
__global__ __launch_bounds__(256) void ocount_kernel(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols,
                                                     const uint32_t* __restrict__ pi, uint64_t nv,
                                                     unsigned long long* __restrict__ cnt, uint32_t* __restrict__ tmp) {
    const uint32_t lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / kWave);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; v < nv; v += nw) {
        const uint32_t x = uniform(pi[v]);
        const uint64_t e00 = uniform(rp[v]), e1 = uniform(rp[v + 1]);
        uint64_t k = 0;
        for (uint64_t e0 = e00; e0 < e1; e0 += kRowUnroll * kWave) {
            uint32_t y[kRowUnroll];
#pragma unroll
            for (int j = 0; j < kRowUnroll; j++) {
                const uint64_t e = e0 + j * kWave + lane;
                y[j] = e < e1 ? cols[e] : 0xffffffffu;
            }
#pragma unroll
            for (int j = 0; j < kRowUnroll; j++) y[j] = y[j] != 0xffffffffu ? pi[y[j]] : 0;
#pragma unroll
            for (int j = 0; j < kRowUnroll; j++) {
                const bool up = y[j] > x;
                const uint64_t bm = ballot(up);
                if (up) tmp[e00 + k + lanes_below(bm)] = y[j];
                k += __popcll(bm);
            }
        }
        if (lane == 0) cnt[x] = k;
    }
}

// N+(x), sorted.  A row of at most 64 ranks: each goes to its rank within
// the row, counted over the lanes; up to kSortMax: bitonic sort in the wave's
// LDS; longer: ranks counted chunk by chunk (ties, which a simple graph has
// none of, by slot).
constexpr uint32_t kSortMax = 1024;
__global__ __launch_bounds__(256) void ofill_kernel(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ pi,
                                                    uint64_t nv, const unsigned long long* __restrict__ orp,
                                                    const uint32_t* __restrict__ tmp, uint32_t* __restrict__ ocols) {
    __shared__ uint32_t sbuf[4][kSortMax];
    uint32_t* buf = sbuf[threadIdx.x / kWave];
    const uint32_t lane = lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / kWave);
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; v < nv; v += nw) {
        const uint32_t x = uniform(pi[v]);
        const uint64_t base = uniform((uint64_t)orp[x]);
        const uint64_t m = uniform((uint64_t)orp[x + 1]) - base;
        if (m == 0) continue;
        const uint32_t* src = tmp + uniform(rp[v]);
        if (m <= kWave) {  // one value per lane: rank by lane reads
            const uint32_t my = lane < m ? src[lane] : 0xffffffffu;
            uint32_t r = 0;
            for (uint32_t q = 0; q < (uint32_t)m; q++) {
                const uint32_t o = read_lane(my, q);
                r += o < my || (o == my && q < lane);
            }
            if (lane < m) ocols[base + r] = my;
            continue;
        }
        if (m <= kSortMax) {
            uint32_t P = 2 * kWave;
            while (P < m) P <<= 1;
            for (uint32_t i = lane; i < P; i += kWave) buf[i] = i < m ? src[i] : 0xffffffffu;
            wave_sync();
            for (uint32_t k = 2; k <= P; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t t = lane; t < P / 2; t += kWave) {
                        const uint32_t i = 2 * t - (t & (j - 1)), l = i + j;  // i has bit j clear
                        const uint32_t a0 = buf[i], a1 = buf[l];
                        if ((a0 > a1) == ((i & k) == 0)) {
                            buf[i] = a1;
                            buf[l] = a0;
                        }
                    }
                    wave_sync();
                }
            }
            for (uint32_t i = lane; i < m; i += kWave) ocols[base + i] = buf[i];
            wave_sync();  // the buffer is refilled by the next row
            continue;
        }
        for (uint64_t t0 = 0; t0 < m; t0 += kWave) {
            const uint64_t i = t0 + lane;
            const uint32_t my = i < m ? src[i] : 0xffffffffu;
            uint64_t r = 0;
            for (uint64_t c0 = 0; c0 < m; c0 += kWave) {
                const uint32_t vc = c0 + lane < m ? src[c0 + lane] : 0xffffffffu;
                const uint32_t nc = (uint32_t)umin64(kWave, m - c0);
                for (uint32_t q = 0; q < nc; q++) {
                    const uint32_t o = read_lane(vc, q);
                    r += o < my || (o == my && c0 + q < i);
                }
            }
            if (i < m) ocols[base + r] = my;
        }
    }
}

// Algorithmic bytes of the counting passes, per pass (the untimed
// accounting run).  A root x reads its two row offsets (16 B); when it has
// at least n-1 members it reads U = N+(x) (4m B) and, for every member U_i
// with a later member, that member's row offsets (16 B) and the cheaper of
// its whole oriented row and one binary search of it per later member:
// 4 min(d_i, a_i (1 + bitlen(d_i))), d_i = |N+(U_i)|, a_i = m-1-i.  The
// root's bytes go to the pass whose LDS rows it fits, by the rules the
// passes apply (pass 1 takes every root of a 2-vertex plan).
__device__ inline int local_pass_of(uint32_t n, uint64_t m) {
    if (m < n - 1 || n == 2) return 0;
    const uint64_t need = local_need(m);
    if (m <= kLU && need <= kLRowWords) return 0;
    if (m <= (uint64_t)kB1U && need <= (uint64_t)kB1RowWords) return 1;
    if (m <= (uint64_t)kB2U && need <= (uint64_t)kB2RowWords) return 2;
    if (m <= (uint64_t)kB3U && need <= (uint64_t)kB3RowWords) return 3;
    return 4;
}

__global__ __launch_bounds__(256) void local_bytes_kernel(const uint64_t* __restrict__ orp,
                                                          const uint32_t* __restrict__ ocols,
                                                          const uint32_t* __restrict__ list, uint64_t nroots,
                                                          uint32_t n, unsigned long long* __restrict__ out) {
    const uint32_t lane = lane_id();
    const uint64_t gwave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    uint64_t acc[5] = {0, 0, 0, 0, 0};
    for (uint64_t t = gwave; t < nroots; t += nw) {
        const uint32_t x = list ? list[t] : (uint32_t)t;
        const uint64_t lo = orp[x], m = orp[x + 1] - lo;
        uint64_t b = 0;
        if (n > 2 && m >= n - 1) {
            for (uint64_t i = lane; i + 1 < m; i += kWave) {
                const uint32_t u = ocols[lo + i];
                const uint64_t d = orp[u + 1] - orp[u], later = m - 1 - i;
                b += 16 + 4 * umin64(d, later * (1 + bitlen(d)));
            }
            b = wave_sum(b);
            b += 4 * m;
        }
        b += 16;
        const int pass = local_pass_of(n, m);
#pragma unroll
        for (int k = 0; k < 5; k++)
            if (k == pass) acc[k] += b;
    }
    if (lane == 0)
        for (int k = 0; k < 5; k++)
            if (acc[k]) atomicAdd(&out[k], (unsigned long long)acc[k]);
}

// Ranks of the sources: the roots of a view whose sources are a subset.
__global__ __launch_bounds__(256) void oroot_kernel(const uint32_t* __restrict__ src, uint64_t ns,
                                                    const uint32_t* __restrict__ pi, uint32_t* __restrict__ roots) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += (uint64_t)gridDim.x * blockDim.x)
        roots[i] = pi[src[i]];
}

// Build pi, orp and ocols of the uploaded view on the context's stream.
static int build_oriented(g2_ctx* c, const uint64_t** orp, const uint32_t** ocols, const uint32_t** rank) {
    const uint64_t nv = c->nv, ne = c->ne;
    // keys (degree << 32 | id): the degree needs bitlen(max degree) bits above the id
    const int nbits = (int)(32 + std::max<uint64_t>(1, 64 - __builtin_clzll(std::max<uint64_t>(c->max_degree, 1))));
    size_t t_sort = 0, t_scan = 0;
    hipError_t e;
    if ((e = hipcub::DeviceRadixSort::SortKeys(nullptr, t_sort, (unsigned long long*)nullptr,
                                                (unsigned long long*)nullptr, (int)nv, 0, nbits, c->stream)))
        return hip_err(c, e, "sort size");
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, (unsigned long long*)nullptr,
                                               (unsigned long long*)nullptr, (int)(nv + 1), c->stream)))
        return hip_err(c, e, "scan size");
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    // layout: keys u64[nv] | sorted u64[nv] | cnt u64[nv+1] | orp u64[nv+1] | pi u32[nv] |
    //         tmp u32[ne] | ocols u32[ne] | temp
    const uint64_t o_sorted = al(nv * 8), o_cnt = o_sorted + al(nv * 8), o_orp = o_cnt + al((nv + 1) * 8),
                   o_pi = o_orp + al((nv + 1) * 8), o_tmp = o_pi + al(nv * 4), o_ocols = o_tmp + al(ne * 4),
                   o_temp = o_ocols + al(std::max<uint64_t>(ne, 1) * 4), total = o_temp + al(std::max(t_sort, t_scan));
    int rc;
    if ((rc = ensure(c, c->orient, total, "oriented rows"))) return rc;
    char* base = (char*)c->orient.ptr;
    // the rows depend on the view only: built once per upload, kept for
    // later runs (any plan) until the next upload
    const bool cached = c->orient_valid;
    auto* keys = (unsigned long long*)base;
    auto* sorted = (unsigned long long*)(base + o_sorted);
    auto* cnt = (unsigned long long*)(base + o_cnt);
    auto* d_orp = (unsigned long long*)(base + o_orp);
    auto* pi = (uint32_t*)(base + o_pi);
    auto* tmp = (uint32_t*)(base + o_tmp);
    auto* d_ocols = (uint32_t*)(base + o_ocols);
    const auto* rp = (const uint64_t*)c->rp.ptr;
    const auto* cols = (const uint32_t*)c->cols.ptr;
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((nv + 255) / 256, 8192));
    const uint64_t wgrid = (uint64_t)c->num_cus * 32 / 4;  // 32 waves per CU: latency-bound row walks
    if (cached) {
        *orp = (const uint64_t*)d_orp;
        *ocols = d_ocols;
        *rank = pi;
        return G2_OK;
    }
    hipLaunchKernelGGL(okey_kernel, dim3(grid), dim3(256), 0, c->stream, (const uint32_t*)c->deg.ptr, nv, keys);
    size_t tb = t_sort;
    if ((e = hipcub::DeviceRadixSort::SortKeys(base + o_temp, tb, keys, sorted, (int)nv, 0, nbits, c->stream)))
        return hip_err(c, e, "rank sort");
    hipLaunchKernelGGL(opi_kernel, dim3(grid), dim3(256), 0, c->stream, sorted, nv, pi);
    if ((e = hipMemsetAsync(cnt + nv, 0, 8, c->stream))) return hip_err(c, e, "memset count");
    hipLaunchKernelGGL(ocount_kernel, dim3(wgrid), dim3(256), 0, c->stream, rp, cols, pi, nv, cnt, tmp);
    tb = t_scan;
    if ((e = hipcub::DeviceScan::ExclusiveSum(base + o_temp, tb, cnt, d_orp, (int)(nv + 1), c->stream)))
        return hip_err(c, e, "oriented scan");
    hipLaunchKernelGGL(ofill_kernel, dim3(wgrid), dim3(256), 0, c->stream, rp, pi, nv, d_orp, tmp, d_ocols);
    if ((e = hipGetLastError())) return hip_err(c, e, "oriented rows kernels");
    *orp = (const uint64_t*)d_orp;
    *ocols = d_ocols;
    *rank = pi;
    c->orient_valid = true;
    return G2_OK;
}

// ---- triangle features (learned scheduler) ---------------------------------
// Every triangle x < y < z (ranks) is found once, at its lowest rank x, as
// the pair (y, z) of U(x) = N+(x) with z in N+(y): the local rows' work list
// without the rows (build_chunk MODE 2).  Its three vertices get one count
// each and its three oriented entries (x,y), (x,z), (y,z) their flags.
__global__ __launch_bounds__(256) void tri_kernel(LocalArgs a) {
    __shared__ uint32_t lU[kLocalBlockWaves][kLU];
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / kWave;
    uint64_t t = 0, tend = 0;
    while (true) {
        if (t == tend) {
            t = fetch_tasks(&a.ctr[LCTR_NEXT], kFetchRun);
            tend = umin64(t + kFetchRun, a.ntasks);
        }
        if (t >= a.ntasks) break;
        const uint32_t x = uniform((uint32_t)t);
        t++;
        const uint64_t lo = uniform(a.orp[x]);
        const uint32_t m = (uint32_t)(uniform(a.orp[x + 1]) - lo);
        if (m < 2) continue;
        const uint32_t* U = a.ocols + lo;
        wave_sync();  // (the previous root's reads of lU are done)
        if (m <= kLU) {
            for (uint32_t i = lane; i < m; i += kWave) lU[wid][i] = U[i];
            wave_sync();
            U = lU[wid];
        }
        uint32_t hits = 0;
        for (uint32_t a0 = 0; a0 + 1 < m; a0 += kWave) build_chunk<2>(a, U, nullptr, m, 0, a0, lo, hits, UBuckets{});
        const uint64_t tot = wave_sum((uint64_t)hits);
        if (lane == 0 && tot) atomicAdd(&a.tri[x], (unsigned long long)tot);
    }
}

// per rank: oriented entries (either end) that lie on a triangle
__global__ __launch_bounds__(256) void sup_kernel(const uint64_t* __restrict__ orp, const uint32_t* __restrict__ ocols,
                                                  const uint8_t* __restrict__ flag, uint64_t nv, uint32_t* sup) {
    const uint64_t nw = (uint64_t)gridDim.x * blockDim.x / kWave;
    for (uint64_t x = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; x < nv; x += nw) {
        uint32_t c = 0;
        for (uint64_t e = orp[x] + lane_id(); e < orp[x + 1]; e += kWave) {
            if (flag[e]) {
                c++;
                atomicAdd(&sup[ocols[e]], 1u);
            }
        }
        c = (uint32_t)wave_sum((uint64_t)c);
        if (lane_id() == 0 && c) atomicAdd(&sup[x], c);
    }
}

int local_triangle_features(g2_ctx* c, unsigned long long* tri, uint32_t* sup, uint8_t* flag, uint64_t flag_bytes,
                            const uint32_t** rank) {
    if (c->nv + 1 > 0x7fffffffull || c->ne >= (1ull << 32))
        return set_err(c, G2_EINVAL, "oriented triangle features need fewer than 2^31-1 vertices and 2^32 entries");
    if (flag_bytes < c->ne) return set_err(c, G2_EINVAL, "triangle flags: buffer too small");
    int rc;
    if ((rc = ensure(c, c->counters, LCTR_N * 8, "counters"))) return rc;
    const uint64_t* orp = nullptr;
    const uint32_t* ocols = nullptr;
    if ((rc = build_oriented(c, &orp, &ocols, rank))) return rc;
    hipError_t e;
    const uint64_t nv = c->nv;
    if ((e = hipMemsetAsync(c->counters.ptr, 0, LCTR_N * 8, c->stream))) return hip_err(c, e, "memset counters");
    if ((e = hipMemsetAsync(tri, 0, std::max<uint64_t>(nv, 1) * 8, c->stream))) return hip_err(c, e, "memset tri");
    if ((e = hipMemsetAsync(sup, 0, std::max<uint64_t>(nv, 1) * 4, c->stream))) return hip_err(c, e, "memset sup");
    if ((e = hipMemsetAsync(flag, 0, std::max<uint64_t>(c->ne, 1), c->stream))) return hip_err(c, e, "memset flags");
    LocalArgs a{};
    a.orp = orp;
    a.ocols = ocols;
    a.ntasks = nv;
    a.ctr = (unsigned long long*)c->counters.ptr;
    a.tri = tri;
    a.flag = flag;
    const uint64_t blocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * waves_per_cu(c, WK_GM_LOCAL) / kLocalBlockWaves);
    hipLaunchKernelGGL(tri_kernel, dim3(blocks), dim3(kLocalBlockWaves * kWave), 0, c->stream, a);
    const uint64_t wgrid = std::max<uint64_t>(1, (uint64_t)c->num_cus * 32 / 4);
    hipLaunchKernelGGL(sup_kernel, dim3(wgrid), dim3(256), 0, c->stream, orp, ocols, (const uint8_t*)flag, nv, sup);
    if ((e = hipGetLastError())) return hip_err(c, e, "triangle feature kernels");
    return G2_OK;
}

void orient_release(g2_ctx* c) {
    if (c->orient.ptr) (void)hipFree(c->orient.ptr);
    c->used -= c->orient.bytes;
    c->orient = DevBuf{};
    c->orient_valid = false;
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
    // clique plans: position i >= 2 adjacent to all of 1..i-1 and above i-1
    // (NON_EQUAL between adjacent positions always holds)
    bool clq = n >= 3;
    for (uint32_t i = 2; i < n; i++)
        clq = clq && L.bn_mask[i] == (1u << i) - 2 && ((L.gt_mask[i] >> (i - 1)) & 1) && !L.lt_mask[i];
    L.clique = clq ? 1 : 0;
    c->lprog = L;
    c->local_ok = true;
}

enum class LocalPass { Wave, Group256, Group512, Group1024 };

template <int N, bool CLQ>
static void launch_local(LocalPass pass, uint64_t blocks, uint32_t bw, hipStream_t s, const LocalArgs& a) {
    if constexpr (N > 2) {
        switch (pass) {
            case LocalPass::Group256:
                hipLaunchKernelGGL((gm_local_block_kernel<N, CLQ, kB1Threads, kB1RowWords, kB1U, kB1Buckets>), dim3(blocks),
                                   dim3(kB1Threads), 0, s, a);
                return;
            case LocalPass::Group512:
                hipLaunchKernelGGL((gm_local_block_kernel<N, CLQ, kB2Threads, kB2RowWords, kB2U, kB2Buckets>), dim3(blocks),
                                   dim3(kB2Threads), 0, s, a);
                return;
            case LocalPass::Group1024:
                hipLaunchKernelGGL((gm_local_block_kernel<N, CLQ, kB3Threads, kB3RowWords, kB3U, kB3Buckets>), dim3(blocks),
                                   dim3(kB3Threads), 0, s, a);
                return;
            default:
                break;
        }
    }
    hipLaunchKernelGGL((gm_local_kernel<N, CLQ>), dim3(blocks), dim3(bw * kWave), 0, s, a);
}

// Passes over the roots: rows in a wave's LDS; in a 256-, 512- and
// 1024-thread workgroup's LDS; in per-wave global arenas sized to the largest
// remaining root.  Each pass takes the roots the previous one deferred; the
// deferred lists alternate between two buffers.
int gm_local_run(g2_ctx* c, uint64_t* count, g2_stats* st, bool bytes) {
    if (c->nv + 1 > 0x7fffffffull) return set_err(c, G2_EINVAL, "local-bitmap matching supports fewer than 2^31-1 vertices");
    const uint32_t bw = std::min<uint32_t>(c->tune.block_waves, kLocalBlockWaves);
    const uint64_t nblocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * waves_per_cu(c, WK_GM_LOCAL) / bw);
    // roots: every rank of a whole-graph view, else the ranks of the sources
    // (a clique is then counted in the view holding its lowest-ranked vertex)
    const bool all = c->sources_all;
    const uint64_t nroots = all ? c->nv : c->ns;
    const uint64_t list_len = std::max<uint64_t>(nroots, 1);
    int rc;
    if ((rc = ensure(c, c->counters, LCTR_N * 8, "counters"))) return rc;
    if ((rc = ensure(c, c->big_list2, list_len * 4, "deferred roots"))) return rc;
    if ((rc = ensure(c, c->list, list_len * 4, "deferred roots"))) return rc;
    if (!all && (rc = ensure(c, c->roots, list_len * 4, "root ranks"))) return rc;
    hipError_t e;
    auto* ctr = (unsigned long long*)c->counters.ptr;
    uint32_t* lists[2] = {(uint32_t*)c->big_list2.ptr, (uint32_t*)c->list.ptr};
    float tot = 0;
    double kernel_ms = 0;
    if ((e = hipEventRecord(c->ev[0], c->stream))) return hip_err(c, e, "event");
    const uint64_t* orp = nullptr;
    const uint32_t* ocols = nullptr;
    const uint32_t* pi = nullptr;
    if ((rc = build_oriented(c, &orp, &ocols, &pi))) return rc;
    if (!all && c->ns) {
        const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((c->ns + 255) / 256, 8192));
        hipLaunchKernelGGL(oroot_kernel, dim3(grid), dim3(256), 0, c->stream, (const uint32_t*)c->src.ptr, c->ns, pi,
                           (uint32_t*)c->roots.ptr);
        if ((e = hipGetLastError())) return hip_err(c, e, "root ranks");
    }
    if ((e = hipMemsetAsync(ctr, 0, LCTR_N * 8, c->stream))) return hip_err(c, e, "memset counters");
    LocalArgs a{};
    a.orp = orp;
    a.ocols = ocols;
    a.prog = c->lprog;
    a.ctr = ctr;
    auto launch = [&](LocalPass pass, uint64_t blocks) {
        const bool clq = c->lprog.clique != 0;
        switch (c->lprog.n) {
            case 2: launch_local<2, false>(pass, blocks, bw, c->stream, a); break;
            case 3: clq ? launch_local<3, true>(pass, blocks, bw, c->stream, a) : launch_local<3, false>(pass, blocks, bw, c->stream, a); break;
            case 4: clq ? launch_local<4, true>(pass, blocks, bw, c->stream, a) : launch_local<4, false>(pass, blocks, bw, c->stream, a); break;
            case 5: clq ? launch_local<5, true>(pass, blocks, bw, c->stream, a) : launch_local<5, false>(pass, blocks, bw, c->stream, a); break;
            case 6: clq ? launch_local<6, true>(pass, blocks, bw, c->stream, a) : launch_local<6, false>(pass, blocks, bw, c->stream, a); break;
            case 7: clq ? launch_local<7, true>(pass, blocks, bw, c->stream, a) : launch_local<7, false>(pass, blocks, bw, c->stream, a); break;
            default: clq ? launch_local<8, true>(pass, blocks, bw, c->stream, a) : launch_local<8, false>(pass, blocks, bw, c->stream, a); break;
        }
        return hipGetLastError();
    };
    unsigned long long h[LCTR_N];
    double pass_ms[5] = {0, 0, 0, 0, 0};
    int pass_no = 0;
    auto timed = [&](LocalPass pass, uint64_t blocks) -> int {
        if ((e = hipMemsetAsync(ctr + LCTR_NEXT, 0, 8, c->stream))) return hip_err(c, e, "memset next");
        if ((e = hipEventRecord(c->ev[1], c->stream))) return hip_err(c, e, "event");
        if ((e = launch(pass, blocks))) return hip_err(c, e, "local launch");
        if ((e = hipEventRecord(c->ev[2], c->stream))) return hip_err(c, e, "event");
        if ((e = hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H counters");
        if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "local sync");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[1], c->ev[2]);
        kernel_ms += ms;
        pass_ms[pass_no] = ms;
        return G2_OK;
    };
    // pass 1: every root, rows in the wave's LDS
    a.list = all ? nullptr : (const uint32_t*)c->roots.ptr;
    a.ntasks = nroots;
    a.scratch = nullptr;
    a.scratch_words = 0;
    a.big_list = lists[0];
    a.defer_ctr = LCTR_DEFER;
    if ((rc = timed(LocalPass::Wave, nblocks))) return rc;
    uint32_t launches = 1;
    const uint64_t n1 = h[LCTR_DEFER];
    uint64_t n = n1, need = h[LCTR_DEFER + 1];
    uint64_t pass_roots[5] = {nroots - n1, 0, 0, 0, 0};
    int in = 0;  // list holding the roots of the next pass
    const struct { LocalPass pass; int per_cu; } groups[3] = {
        {LocalPass::Group256, kB1PerCu}, {LocalPass::Group512, kB2PerCu}, {LocalPass::Group1024, kB3PerCu}};
    for (int g = 0; g < 3 && n; g++) {  // workgroup passes
        pass_no = g + 1;
        a.list = lists[in];
        a.ntasks = n;
        a.big_list = lists[in ^ 1];
        a.defer_ctr = LCTR_DEFER + 2 * (g + 1);
        if ((rc = timed(groups[g].pass, std::min<uint64_t>((uint64_t)c->num_cus * groups[g].per_cu, n)))) return rc;
        launches++;
        pass_roots[g + 1] = n - h[a.defer_ctr];
        n = h[a.defer_ctr];
        need = h[a.defer_ctr + 1];
        in ^= 1;
    }
    if (n) {  // last pass: per-wave global arenas
        pass_no = 4;
        uint64_t bwaves = std::min<uint64_t>(std::max<uint32_t>(c->tune.big_waves, 1), n);
        while (bwaves > 1 && c->budget && c->used - c->big_scratch.bytes + bwaves * need * 8 > c->budget) bwaves /= 2;
        bwaves = (bwaves + bw - 1) / bw * bw;
        if ((rc = ensure(c, c->big_scratch, bwaves * need * 8, "large-root arenas"))) return rc;
        a.list = lists[in];
        a.ntasks = n;
        a.scratch = (uint64_t*)c->big_scratch.ptr;
        a.scratch_words = need;
        a.big_list = nullptr;
        if ((rc = timed(LocalPass::Wave, bwaves / bw))) return rc;
        launches++;
        pass_roots[4] = n;
    }
    if ((e = hipEventRecord(c->ev[3], c->stream))) return hip_err(c, e, "event");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "local sync");
    (void)hipEventElapsedTime(&tot, c->ev[0], c->ev[3]);
    if (h[LCTR_ERR]) return set_err(c, G2_EBUDGET, "a root did not fit the large-root arena");
    unsigned long long pass_bytes[5] = {0, 0, 0, 0, 0};
    if (bytes) {  // accounting (untimed): the oriented rows each pass had to read
        if ((e = hipMemsetAsync(ctr + LCTR_DEFER, 0, 5 * 8, c->stream))) return hip_err(c, e, "memset bytes");
        const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((nroots + 3) / 4, (uint64_t)c->num_cus * 32));
        hipLaunchKernelGGL(local_bytes_kernel, dim3(grid), dim3(256), 0, c->stream, orp, ocols,
                           all ? nullptr : (const uint32_t*)c->roots.ptr, nroots, c->lprog.n, ctr + LCTR_DEFER);
        if ((e = hipGetLastError())) return hip_err(c, e, "local bytes launch");
        if ((e = hipMemcpyAsync(pass_bytes, ctr + LCTR_DEFER, sizeof(pass_bytes), hipMemcpyDeviceToHost, c->stream)))
            return hip_err(c, e, "D2H bytes");
        if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "local bytes sync");
    }
    if (count) *count = h[LCTR_COUNT];
    if (st) {
        st->kernel_ms = kernel_ms;
        st->launches = launches;
        st->total_ms = tot;
        st->nodes = h[LCTR_NODES];
        st->alg_bytes = 0;
        for (int k = 0; k < 5; k++) {
            st->alg_bytes += pass_bytes[k];
            st->local_pass_bytes[k] = pass_bytes[k];
            st->local_pass_ms[k] = pass_ms[k];
        }
        st->big_roots = n1;
        st->max_scratch_words = need;
        st->wave_idle_frac = 0;
        st->donations = 0;
        st->local_bitmap = 1;
        for (int k = 0; k < 5; k++) st->local_pass_roots[k] = pass_roots[k];
    }
    return G2_OK;
}

}  // namespace g2
