// ============================================================================
// oracle/oracle.cpp -- CPU restatement of G2-AIMD's subgraph-search hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: it may be
// loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
// leg, and by nothing else.  The product (libg2hip.so, the bk/run CLIs) never
// links, loads or calls it.
//
// What it restates (reference file:line):
//   * orc_mce  -- pivoted Bron-Kerbosch exactly as the reference's device code
//     walks it, on labelled lists R/P/X/Q:
//       root node            app_BK/BK.h:323-355   ([v:r] ++ [u: x if u<v else p])
//       clique test          app_BK/BK.h:78-91     (no P and no X)
//       pivot                app_BK/BK.h:282-321   (argmax |P & N(u)| over P u X,
//                                                   ties -> last in list order)
//       Q = P \ N(pivot)     app_BK/BK.h:445-464
//       children in order    app_BK/BK.h:467-473, 236-280 (Q-labelled survivors
//                                                   become x if v<q else p)
//     The reference walks the tree breadth-first over a global queue; the tree
//     itself is a deterministic function of the graph, so this depth-first walk
//     visits the same node set.  Per-root node counts are therefore exact.
//   * orc_gm   -- plan-driven extension of partial matches exactly as
//     app_gmatch/gmatch.h:83-674 (non-shared path :510-671: parent = back
//     neighbour with the smallest row, first on ties; shared path :273-509:
//     pre-intersection once per prefix, after-filter per candidate).
//   * binarySearch (system/util.h:19-30) is restated as a membership test on a
//     sorted row.  The reference reads arr[low] on an EMPTY range (a latent
//     out-of-range read); here an empty range is simply "not a member".  Empty
//     rows are only reachable in views built with too small a hop, where the
//     reference result is undefined.
//
// Algorithmic byte model (the "per-unit figure" the roofline uses):
//   c(A,u) = 4*min(|A| + d(u), |A|*(1 + ceil(log2(d(u)+1))))
//   BK node:  5s + sum_{u in P u X} c(P,u) + c(P,u*) + sum_{q in Q}[c(PuX,q) + 5*|child_q|]
//             (|child_q| counts every entry written for the child, q included;
//              a maximal clique costs 5s)
//   GM node (extension of a match of length i through parent row pi):
//             4i + 4d(pi) + sum_{cand passing conds} sum_{b in B'} c({cand},b) + W,
//             W = 4(i+1)|out| (EXPAND) or 4(i+|out|) (PREFIX), 0 at the last level;
//             shared positions: the pre-intersection (4i + 4d(pi_pre) + membership
//             tests) is charged once per prefix node, the after-filter
//             c(pre, after) + W once per candidate.
//
// Parity pins (tests/): brute-force maximal cliques by subset enumeration
// (orc_bruteforce_mce, n <= 20), closed forms (K_n, K_{a,b}, Moon-Moser,
// cycles, Turan graphs, disjoint unions), brute-force injective
// homomorphisms / |Aut| (orc_bruteforce_pattern), the reference's own plan
// tables compiled from its sources (oracle/_ref), and README.md:65-66 style
// clique counts by an independent node-iterator (orc_kclique_count); per-root
// node counts and modelled bytes worked by hand on small graphs.  The tests
// also plant mistakes in copies of this file and require the pins to catch
// each one.
// ============================================================================
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr char LR = 'r', LP = 'p', LX = 'x', LQ = 'q';

struct Csr {
    const uint64_t* rp;
    const uint32_t* cols;
    uint64_t nv;
    uint64_t deg(uint32_t v) const { return rp[v + 1] - rp[v]; }
};

// membership of val in the sorted row cols[lo, hi)  (system/util.h:19-30)
inline bool member(const uint32_t* cols, uint64_t lo, uint64_t hi, uint32_t val) {
    if (lo >= hi) return false;
    uint64_t h = hi - 1;
    while (lo < h) {
        uint64_t mid = lo + (h - lo) / 2;
        uint32_t a = cols[mid];
        if (a == val) return true;
        if (a < val) lo = mid + 1; else h = mid;
    }
    return cols[lo] == val;
}

inline uint64_t bitlen(uint64_t d) { return d == 0 ? 0 : 64 - __builtin_clzll(d); }

// c(A,u) of the byte model, |A| = a, d(u) = d
inline uint64_t cost_isect(uint64_t a, uint64_t d) {
    uint64_t merge = a + d;
    uint64_t bs = a * (1 + bitlen(d));
    return 4 * std::min(merge, bs);
}

struct Ent { uint32_t v; char l; };

struct MceAcc {
    uint64_t nodes = 0, cliques = 0, bytes = 0;
    std::vector<uint64_t> hist;
};

// one BK node (app_BK/BK.h:405-443 for one dequeued subgraph), depth first
void bk_node(const Csr& g, std::vector<Ent>& node, uint32_t nR, MceAcc& acc) {
    acc.nodes++;
    const uint64_t s = node.size();
    acc.bytes += 5 * s;
    bool anyPX = false;
    for (auto& e : node) if (e.l == LP || e.l == LX) { anyPX = true; break; }
    if (!anyPX) {  // examineClique (BK.h:78-91) -> maximal clique of size |R|
        acc.cliques++;
        uint64_t b = std::min<uint64_t>(nR, acc.hist.size() - 1);
        acc.hist[b]++;
        return;
    }
    // selectPivot (BK.h:282-321)
    std::vector<uint32_t> pl;
    for (auto& e : node) if (e.l == LP) pl.push_back(e.v);
    uint32_t maxc = 0, pivot = 0;
    uint64_t nPX = 0;
    for (auto& e : node) {
        if (e.l == LR) continue;
        nPX++;
        const uint32_t u = e.v;
        const uint64_t lo = g.rp[u], hi = g.rp[u + 1];
        uint32_t cnt = 0;
        for (uint32_t p : pl) cnt += member(g.cols, lo, hi, p);
        acc.bytes += cost_isect(pl.size(), hi - lo);
        if (cnt >= maxc) { maxc = cnt; pivot = u; }
    }
    // expandClique (BK.h:445-474): relabel P \ N(pivot) as Q
    const uint64_t plo = g.rp[pivot], phi = g.rp[pivot + 1];
    acc.bytes += cost_isect(pl.size(), phi - plo);
    for (auto& e : node)
        if (e.l == LP && !member(g.cols, plo, phi, e.v)) e.l = LQ;
    // one child per Q vertex, in list order (BK.h:467-473, 236-280)
    std::vector<Ent> child;
    for (size_t i = 0; i < node.size(); i++) {
        if (node[i].l != LQ) continue;
        const uint32_t q = node[i].v;
        const uint64_t qlo = g.rp[q], qhi = g.rp[q + 1];
        child.clear();
        child.push_back({q, LR});
        for (auto& e : node) {
            if (e.l == LR) { child.push_back(e); continue; }
            if (!member(g.cols, qlo, qhi, e.v)) continue;
            char l = e.l;
            if (l == LQ) l = (e.v < q) ? LX : LP;
            child.push_back({e.v, l});
        }
        acc.bytes += cost_isect(nPX, qhi - qlo) + 5 * child.size();
        bk_node(g, child, nR + 1, acc);
    }
}

}  // namespace

extern "C" {

// Maximal clique enumeration over the given roots.  Returns 0.
//   hist[k] = number of maximal cliques with k vertices (k >= hist_len-1 land in
//   the last bin).  per_root_nodes / per_root_bytes (nullable, length ns) get the
//   search-tree node count and modelled bytes of each root's subtree.
int orc_mce(const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint64_t ne,
            const uint32_t* src, uint64_t ns, int nthreads,
            uint64_t* total, uint64_t* hist, uint32_t hist_len,
            uint64_t* per_root_nodes, uint64_t* per_root_bytes,
            uint64_t* total_nodes, uint64_t* total_bytes) {
    (void)ne;
    if (hist_len < 2) return -1;
    Csr g{rp, cols, nv};
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<uint64_t> ghist(hist_len, 0);
    uint64_t gcl = 0, gnodes = 0, gbytes = 0;
#pragma omp parallel num_threads(nthreads)
    {
        MceAcc acc;
        acc.hist.assign(hist_len, 0);
        std::vector<Ent> root;
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)ns; i++) {
            const uint32_t v = src[i];
            const uint64_t lo = rp[v], hi = rp[v + 1];
            uint64_t n0 = acc.nodes, b0 = acc.bytes;
            if (hi > lo) {  // deg-0 roots are skipped (BK.h:335-336)
                root.clear();
                root.push_back({v, LR});
                for (uint64_t j = lo; j < hi; j++) {
                    uint32_t u = cols[j];
                    root.push_back({u, u < v ? LX : LP});
                }
                bk_node(g, root, 1, acc);
            }
            if (per_root_nodes) per_root_nodes[i] = acc.nodes - n0;
            if (per_root_bytes) per_root_bytes[i] = acc.bytes - b0;
        }
#pragma omp critical
        {
            for (uint32_t k = 0; k < hist_len; k++) ghist[k] += acc.hist[k];
            gcl += acc.cliques;
            gnodes += acc.nodes;
            gbytes += acc.bytes;
        }
    }
    if (total) *total = gcl;
    if (hist) std::memcpy(hist, ghist.data(), hist_len * sizeof(uint64_t));
    if (total_nodes) *total_nodes = gnodes;
    if (total_bytes) *total_bytes = gbytes;
    return 0;
}

// Plan layout shared with the product's g2_plan (include/g2hip.h); the tests
// check that both ctypes mirrors have the same size and offsets.
#define ORC_MAXQ 8
struct orc_plan {
    uint32_t n;
    uint32_t root_degree;
    int32_t hop;
    uint32_t seq[ORC_MAXQ];                     // position -> query vertex
    uint32_t rev[ORC_MAXQ];                     // query vertex -> position
    uint32_t bn_count[ORC_MAXQ];                // by position
    uint32_t bn[ORC_MAXQ][ORC_MAXQ];            // by position: query vertices
    uint32_t cond_count[ORC_MAXQ];              // by query vertex
    uint32_t cond[ORC_MAXQ][ORC_MAXQ][2];       // by query vertex: (op, query vertex)
    uint32_t shared[ORC_MAXQ];                  // by position
    uint32_t pre_bn_count[ORC_MAXQ];            // by position
    uint32_t pre_bn[ORC_MAXQ][ORC_MAXQ];
    uint32_t after_bn_count[ORC_MAXQ];
    uint32_t after_bn[ORC_MAXQ][ORC_MAXQ];
    uint32_t pre_cond_count[ORC_MAXQ];          // by query vertex
    uint32_t pre_cond[ORC_MAXQ][ORC_MAXQ][2];
    uint32_t after_cond_count[ORC_MAXQ];
    uint32_t after_cond[ORC_MAXQ][ORC_MAXQ][2];
    uint32_t strategy[ORC_MAXQ + 1];            // by position: 0 EXPAND 1 PREFIX 2 COUNT
    uint32_t moving_lvl[ORC_MAXQ];
};

}  // extern "C"

namespace {

enum { C_LESS = 0, C_LARGER = 1, C_NEQ = 2 };

struct GmCtx {
    Csr g;
    const orc_plan* p;
    uint64_t count = 0, bytes = 0;
    uint32_t M[ORC_MAXQ];  // matched data vertex by position
};

// condition check of candidate w for query vertex u against matched vertices
// (gmatch.h:559-588): "op j" on u means  M(u) op M(j)
inline bool conds_ok(const GmCtx& c, uint32_t cnt, const uint32_t (*cond)[2], uint32_t w) {
    for (uint32_t k = 0; k < cnt; k++) {
        const uint32_t op = cond[k][0];
        const uint32_t m = c.M[c.p->rev[cond[k][1]]];
        if (op == C_LESS) { if (m <= w) return false; }
        else if (op == C_LARGER) { if (m >= w) return false; }
        else { if (m == w) return false; }
    }
    return true;
}

// pick the back neighbour whose matched row is smallest, first on ties
inline uint32_t min_row(const GmCtx& c, uint32_t cnt, const uint32_t* bns) {
    uint32_t best = bns[0];
    uint64_t bd = c.g.deg(c.M[c.p->rev[bns[0]]]);
    for (uint32_t i = 1; i < cnt; i++) {
        uint64_t d = c.g.deg(c.M[c.p->rev[bns[i]]]);
        if (d < bd) { bd = d; best = bns[i]; }
    }
    return best;
}

// filter the parent's row by conditions and membership in the other rows
void filter_row(GmCtx& c, uint32_t cnt, const uint32_t* bns, uint32_t parent,
                uint32_t ccnt, const uint32_t (*cond)[2], std::vector<uint32_t>& out) {
    const uint32_t pm = c.M[c.p->rev[parent]];
    uint64_t bcost = 0;
    for (uint32_t j = 0; j < cnt; j++)
        if (bns[j] != parent) bcost += cost_isect(1, c.g.deg(c.M[c.p->rev[bns[j]]]));
    c.bytes += 4 * c.g.deg(pm);
    for (uint64_t e = c.g.rp[pm]; e < c.g.rp[pm + 1]; e++) {
        const uint32_t w = c.g.cols[e];
        if (!conds_ok(c, ccnt, cond, w)) continue;
        c.bytes += bcost;
        bool ok = true;
        for (uint32_t j = 0; j < cnt && ok; j++) {
            if (bns[j] == parent) continue;
            const uint32_t bm = c.M[c.p->rev[bns[j]]];
            ok = member(c.g.cols, c.g.rp[bm], c.g.rp[bm + 1], w);
        }
        if (ok) out.push_back(w);
    }
}

inline uint64_t store_bytes(const GmCtx& c, uint32_t id, uint64_t nout) {
    const uint32_t n = c.p->n;
    if (id + 1 == n) return 0;  // counted, not stored
    if (c.p->strategy[id + 1] == 1) return 4 * (id + nout);      // PREFIX
    return 4 * (uint64_t)(id + 1) * nout;                        // EXPAND
}

void gm_extend(GmCtx& c, uint32_t id);

// candidates for a shared position id given the matched prefix and the
// candidate list of position id-1 (gmatch.h:273-509)
void gm_shared(GmCtx& c, uint32_t id, const std::vector<uint32_t>& prev_cands) {
    const orc_plan* p = c.p;
    const uint32_t u = p->seq[id];
    // pre-intersection, once per prefix node
    std::vector<uint32_t> pre;
    const uint32_t parent = min_row(c, p->pre_bn_count[id], p->pre_bn[id]);
    c.bytes += 4 * (id - 1);
    filter_row(c, p->pre_bn_count[id], p->pre_bn[id], parent,
               p->pre_cond_count[u], p->pre_cond[u], pre);
    std::vector<uint32_t> cands;
    for (uint32_t cv : prev_cands) {
        c.M[id - 1] = cv;
        cands.clear();
        uint64_t after_cost = 0;
        for (uint32_t j = 0; j < p->after_bn_count[id]; j++)
            after_cost += cost_isect(pre.size(), c.g.deg(c.M[p->rev[p->after_bn[id][j]]]));
        c.bytes += after_cost;
        for (uint32_t w : pre) {
            if (!conds_ok(c, p->after_cond_count[u], p->after_cond[u], w)) continue;
            bool ok = true;
            for (uint32_t j = 0; j < p->after_bn_count[id] && ok; j++) {
                const uint32_t bm = c.M[p->rev[p->after_bn[id][j]]];
                ok = member(c.g.cols, c.g.rp[bm], c.g.rp[bm + 1], w);
            }
            if (ok) cands.push_back(w);
        }
        c.bytes += store_bytes(c, id, cands.size());
        if (id + 1 == p->n) { c.count += cands.size(); continue; }
        if (p->shared[id + 1]) { gm_shared(c, id + 1, cands); continue; }
        for (uint32_t w : cands) { c.M[id] = w; gm_extend(c, id + 1); }
    }
}

// extend a match of length id by position id through the non-shared path
// (gmatch.h:510-671)
void gm_extend(GmCtx& c, uint32_t id) {
    const orc_plan* p = c.p;
    const uint32_t u = p->seq[id];
    const uint32_t parent = min_row(c, p->bn_count[id], p->bn[id]);
    std::vector<uint32_t> cands;
    c.bytes += 4 * id;
    filter_row(c, p->bn_count[id], p->bn[id], parent, p->cond_count[u], p->cond[u], cands);
    c.bytes += store_bytes(c, id, cands.size());
    if (id + 1 == p->n) { c.count += cands.size(); return; }
    if (p->shared[id + 1]) { gm_shared(c, id + 1, cands); return; }
    for (uint32_t w : cands) { c.M[id] = w; gm_extend(c, id + 1); }
}

}  // namespace

extern "C" {

// Match count of one plan over the given roots (u64, overflow-free semantics).
int orc_gm(const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint64_t ne,
           const uint32_t* src, uint64_t ns, const orc_plan* plan, int nthreads,
           uint64_t* count, uint64_t* per_root_count, uint64_t* total_bytes) {
    (void)ne;
    if (!plan || plan->n < 2 || plan->n > ORC_MAXQ) return -1;
    if (plan->shared[1]) return -2;  // position 1 always extends from the root
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    uint64_t gcount = 0, gbytes = 0;
#pragma omp parallel num_threads(nthreads)
    {
        GmCtx c{Csr{rp, cols, nv}, plan};
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)ns; i++) {
            uint64_t c0 = c.count;
            c.M[0] = src[i];
            c.bytes += 4;  // the 1-vertex partial match (gmatch.h:67-81)
            gm_extend(c, 1);
            if (per_root_count) per_root_count[i] = c.count - c0;
        }
#pragma omp critical
        {
            gcount += c.count;
            gbytes += c.bytes;
        }
    }
    if (count) *count = gcount;
    if (total_bytes) *total_bytes = gbytes;
    return 0;
}

// ---------------------------------------------------------------------------
// Distinct 2-hop ball size per source, the cost model's "2-hop frontier"
// input: a HyperLogLog estimate (Flajolet, Fusy, Gandouet, Meunier 2007) of
// |B2(v)| = |{v} u N(v) u N(N(v))|, with m = 64 registers, the splitmix64
// finaliser as hash, register index = top 6 bits, rank = leading zeros of the
// other 58 bits + 1 capped at 32.  Each vertex u gets the registers of its
// closed row N[u] = N(u) u {u}; a source's registers are the register-wise
// maximum over N[v] and every N[u], u in N(v).  Estimate: E = alpha m^2 /
// sum 2^-reg (alpha_64 = 0.709), evaluated as alpha m^2 2^32 / sum 2^(32-reg)
// with an exact integer sum; when E <= 5m/2 and V registers are zero (V > 0),
// linear counting m ln(m/V) instead.  Output: (int64)(E + 0.5).  No reference
// counterpart (the reference has no root features); the device kernel
// computes the same integers and doubles.
static inline uint64_t hll_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline void hll_add(uint8_t* reg, uint32_t w) {
    const uint64_t h = hll_mix(w);
    const uint32_t idx = (uint32_t)(h >> 58);
    const uint64_t rest = h << 6;
    const uint32_t rho = rest ? std::min<uint32_t>((uint32_t)__builtin_clzll(rest) + 1, 32u) : 32u;
    reg[idx] = std::max<uint8_t>(reg[idx], (uint8_t)rho);
}
static inline int64_t hll_estimate(const uint8_t* reg) {
    uint64_t sum = 0;
    uint32_t zeros = 0;
    for (int l = 0; l < 64; l++) {
        sum += 1ull << (32 - reg[l]);
        zeros += reg[l] == 0;
    }
    double e = (0.709 * 4096.0 * 4294967296.0) / (double)sum;
    if (e <= 160.0 && zeros > 0) e = 64.0 * std::log(64.0 / (double)zeros);
    return (int64_t)(e + 0.5);
}

int orc_two_hop(const uint64_t* rp, uint64_t nv, const uint32_t* cols, const uint32_t* sources, uint64_t ns,
                int64_t* out) {
    for (uint64_t i = 0; i < ns; i++)
        if (sources[i] >= nv) return -2;
    std::vector<uint8_t> sk(nv * 64, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < (int64_t)nv; u++) {
        uint8_t* r = &sk[(uint64_t)u * 64];
        hll_add(r, (uint32_t)u);
        for (uint64_t e = rp[u]; e < rp[u + 1]; e++) hll_add(r, cols[e]);
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)ns; i++) {
        const uint32_t v = sources[i];
        uint8_t r[64];
        std::memcpy(r, &sk[(uint64_t)v * 64], 64);
        for (uint64_t e = rp[v]; e < rp[v + 1]; e++) {
            const uint8_t* q = &sk[(uint64_t)cols[e] * 64];
            for (int l = 0; l < 64; l++) r[l] = std::max(r[l], q[l]);
        }
        out[i] = hll_estimate(r);
    }
    return 0;
}

// Independent checkers (brute force), used to pin the restatements above.
// ---------------------------------------------------------------------------

// All maximal cliques of a graph with n <= 24 vertices by subset enumeration
// over adjacency bitmasks.  Isolated vertices are not counted (the reference
// never roots a search at a degree-0 vertex).  hist as in orc_mce.
int orc_bruteforce_mce(const uint64_t* rp, uint64_t nv, const uint32_t* cols,
                       uint64_t* total, uint64_t* hist, uint32_t hist_len) {
    if (nv > 24 || hist_len < 2) return -1;
    const uint32_t n = (uint32_t)nv;
    std::vector<uint32_t> adj(n, 0);
    for (uint32_t v = 0; v < n; v++)
        for (uint64_t e = rp[v]; e < rp[v + 1]; e++) adj[v] |= 1u << cols[e];
    std::memset(hist, 0, hist_len * sizeof(uint64_t));
    uint64_t cnt = 0;
    const uint32_t full = (n == 32) ? 0xffffffffu : ((1u << n) - 1);
    for (uint32_t S = 1; S <= full && S != 0; S++) {
        // clique?
        bool ok = true;
        uint32_t common = full;
        for (uint32_t v = 0; v < n && ok; v++) {
            if (!(S >> v & 1)) continue;
            if ((adj[v] & S) != (S & ~(1u << v))) ok = false;
            common &= adj[v];
        }
        if (!ok) continue;
        if (common & ~S) continue;         // extendable -> not maximal
        uint32_t k = __builtin_popcount(S);
        if (k == 1) continue;              // isolated vertex
        cnt++;
        hist[std::min<uint32_t>(k, hist_len - 1)]++;
    }
    *total = cnt;
    return 0;
}

// Number of injective homomorphisms of the pattern (qn vertices, adjacency
// bitmask per vertex) into the data graph, by plain backtracking with no
// symmetry breaking.  Divide by |Aut(pattern)| (orc_automorphism_count) to
// get the number of distinct (non-induced) occurrences.
int orc_bruteforce_pattern(const uint64_t* rp, uint64_t nv, const uint32_t* cols,
                           uint32_t qn, const uint32_t* qadj, int nthreads, uint64_t* inj_homs) {
    if (qn < 1 || qn > ORC_MAXQ) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    uint64_t tot = 0;
#pragma omp parallel num_threads(nthreads) reduction(+:tot)
    {
        uint32_t M[ORC_MAXQ];
        // recursive lambda via explicit function object
        struct Rec {
            const uint64_t* rp; const uint32_t* cols; uint32_t qn; const uint32_t* qadj;
            uint32_t* M; uint64_t cnt = 0;
            void go(uint32_t i) {
                if (i == qn) { cnt++; return; }
                // candidates: neighbours of some earlier adjacent vertex, or all vertices
                int anchor = -1;
                for (uint32_t j = 0; j < i; j++) if (qadj[i] >> j & 1) { anchor = (int)j; break; }
                if (anchor < 0) return;  // caller orders the pattern BFS-connected
                const uint32_t a = M[anchor];
                for (uint64_t e = rp[a]; e < rp[a + 1]; e++) {
                    const uint32_t w = cols[e];
                    bool ok = true;
                    for (uint32_t j = 0; j < i && ok; j++) {
                        if (M[j] == w) { ok = false; break; }
                        if (qadj[i] >> j & 1) ok = member(cols, rp[M[j]], rp[M[j] + 1], w);
                    }
                    if (!ok) continue;
                    M[i] = w;
                    go(i + 1);
                }
            }
        } rec{rp, cols, qn, qadj, M};
#pragma omp for schedule(dynamic, 64)
        for (int64_t v = 0; v < (int64_t)nv; v++) {
            M[0] = (uint32_t)v;
            rec.go(1);
        }
        tot += rec.cnt;
    }
    *inj_homs = tot;
    return 0;
}

// |Aut| of a pattern by checking every permutation (qn <= 8).
uint64_t orc_automorphism_count(uint32_t qn, const uint32_t* qadj) {
    std::vector<uint32_t> perm(qn);
    for (uint32_t i = 0; i < qn; i++) perm[i] = i;
    uint64_t cnt = 0;
    do {
        bool ok = true;
        for (uint32_t a = 0; a < qn && ok; a++)
            for (uint32_t b = 0; b < qn && ok; b++)
                if (((qadj[a] >> b) & 1) != ((qadj[perm[a]] >> perm[b]) & 1)) ok = false;
        cnt += ok;
    } while (std::next_permutation(perm.begin(), perm.end()));
    return cnt;
}

// k-clique count by an independent method (ordered node-iterator over
// forward neighbourhoods with sorted-merge intersections), k in [3,5].
int orc_kclique_count(const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint32_t k,
                      int nthreads, uint64_t* out) {
    if (k < 3 || k > 5) return -1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    uint64_t tot = 0;
#pragma omp parallel num_threads(nthreads) reduction(+:tot)
    {
        std::vector<uint32_t> a, b, c2;
        auto fwd = [&](uint32_t v, std::vector<uint32_t>& o) {
            o.clear();
            for (uint64_t e = rp[v]; e < rp[v + 1]; e++) if (cols[e] > v) o.push_back(cols[e]);
        };
        auto isect = [&](const std::vector<uint32_t>& x, uint32_t v, std::vector<uint32_t>& o) {
            o.clear();
            uint64_t e = rp[v], ee = rp[v + 1];
            size_t i = 0;
            while (i < x.size() && e < ee) {
                if (x[i] < cols[e]) i++;
                else if (x[i] > cols[e]) e++;
                else { o.push_back(x[i]); i++; e++; }
            }
        };
#pragma omp for schedule(dynamic, 64)
        for (int64_t v = 0; v < (int64_t)nv; v++) {
            fwd((uint32_t)v, a);
            for (uint32_t w : a) {
                isect(a, w, b);
                std::vector<uint32_t> bb;
                for (uint32_t x : b) if (x > w) bb.push_back(x);
                if (k == 3) { tot += bb.size(); continue; }
                for (uint32_t x : bb) {
                    isect(bb, x, c2);
                    uint64_t cnt = 0;
                    std::vector<uint32_t> cc;
                    for (uint32_t y : c2) if (y > x) cc.push_back(y);
                    if (k == 4) { tot += cc.size(); continue; }
                    for (uint32_t y : cc) {
                        for (uint32_t z : cc) if (z > y && member(cols, rp[y], rp[y + 1], z)) cnt++;
                    }
                    tot += cnt;
                }
            }
        }
    }
    *out = tot;
    return 0;
}

// k-cliques whose lowest vertex under the key (degree, id) is one of the given
// sources, k in [2,8]: how a view bin's count is split off the whole-graph
// total when each clique is counted at its lowest-degree vertex (ties to the
// lower id) instead of at its lowest id (app_gmatch/gmatch.h:572-578 with the
// plan's conditions).  Over sources that partition the vertices the counts
// add up to orc_kclique_count's total.  Degrees are the rows' lengths in the
// CSR given, as the view holds them.
int orc_kclique_count_ranked(const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint32_t k,
                             const uint32_t* sources, uint64_t ns, int nthreads, uint64_t* out) {
    if (k < 2 || k > 8) return -1;
    for (uint64_t i = 0; i < ns; i++)
        if (sources[i] >= nv) return -2;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    auto above = [&](uint32_t a, uint32_t b) {  // key(b) > key(a)
        const uint64_t da = rp[a + 1] - rp[a], db = rp[b + 1] - rp[b];
        return db > da || (db == da && b > a);
    };
    uint64_t tot = 0;
#pragma omp parallel num_threads(nthreads) reduction(+:tot)
    {
        std::vector<std::vector<uint32_t>> lvl(k + 1);
        // r-cliques inside the sorted candidate set lvl[d], each taken in id order
        auto count = [&](auto&& self, uint32_t d, uint32_t r) -> uint64_t {
            const std::vector<uint32_t>& C = lvl[d];
            if (r == 1) return C.size();
            uint64_t n = 0;
            for (size_t i = 0; i < C.size(); i++) {
                std::vector<uint32_t>& nx = lvl[d + 1];
                nx.clear();
                const uint32_t c = C[i];
                for (size_t j = i + 1; j < C.size(); j++)
                    if (member(cols, rp[c], rp[c + 1], C[j])) nx.push_back(C[j]);
                n += self(self, d + 1, r - 1);
            }
            return n;
        };
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)ns; i++) {
            const uint32_t s = sources[i];
            lvl[0].clear();
            for (uint64_t e = rp[s]; e < rp[s + 1]; e++)
                if (above(s, cols[e])) lvl[0].push_back(cols[e]);
            tot += count(count, 0, k - 1);
        }
    }
    *out = tot;
    return 0;
}

// Algorithmic bytes of the local-bitmap matching passes (the byte model of
// the kernels that count whole-graph totals of k-clique plans, not of the
// reference's traversal): ranks by (degree, id), U(x) = later-ranked
// neighbours of root x, m = |U(x)|, d_i = number of later-ranked neighbours
// of the i-th member of U(x) in rank order, a_i = m-1-i.  A root costs 16 B
// (its row offsets); with n > 2 and m >= n-1 also 4m (reading U) and, per
// member with a later member, 16 + 4 min(d_i, a_i (1 + bitlen(d_i))): its
// row offsets and the cheaper of its whole row or one binary search of it per
// later member.  per_pass[k] splits the total by the pass whose LDS rows fit
// the root (|U| <= 256 and m * stride <= 448 words; <= 448 and 3136;
// <= 700 and 7700; <= 1040 and 17664; larger), stride = ceil(m/64) | 1,
// with every root of a 2-vertex plan and every root with m < n-1 in the
// first.
int orc_local_bytes(const uint64_t* rp, uint64_t nv, const uint32_t* cols, const uint32_t* sources, uint64_t ns,
                    uint32_t n, uint64_t* per_pass) {
    if (n < 2 || n > 8) return -1;
    for (uint64_t i = 0; i < ns; i++)
        if (sources[i] >= nv) return -2;
    auto deg = [&](uint32_t v) { return rp[v + 1] - rp[v]; };
    auto later = [&](uint32_t a, uint32_t b) {  // key(b) > key(a)
        return deg(b) > deg(a) || (deg(b) == deg(a) && b > a);
    };
    std::vector<uint64_t> up(nv, 0);  // later-ranked neighbours of each vertex
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)nv; v++)
        for (uint64_t e = rp[v]; e < rp[v + 1]; e++) up[v] += later((uint32_t)v, cols[e]);
    uint64_t acc[5] = {0, 0, 0, 0, 0};
    std::vector<uint32_t> U;
    for (uint64_t t = 0; t < ns; t++) {
        const uint32_t x = sources[t];
        U.clear();
        for (uint64_t e = rp[x]; e < rp[x + 1]; e++)
            if (later(x, cols[e])) U.push_back(cols[e]);
        std::sort(U.begin(), U.end(), [&](uint32_t a, uint32_t b) { return later(a, b); });
        const uint64_t m = U.size();
        uint64_t b = 16;
        if (n > 2 && m >= n - 1) {
            b += 4 * m;
            for (uint64_t i = 0; i + 1 < m; i++) {
                const uint64_t d = up[U[i]], a = m - 1 - i;
                b += 16 + 4 * std::min(d, a * (1 + bitlen(d)));
            }
        }
        int pass = 0;
        if (n > 2 && m >= n - 1) {
            const uint64_t need = m * (((m + 63) / 64) | 1);
            if (m <= 256 && need <= 448) pass = 0;
            else if (m <= 448 && need <= 3136) pass = 1;
            else if (m <= 700 && need <= 7700) pass = 2;
            else if (m <= 1040 && need <= 17664) pass = 3;
            else pass = 4;
        }
        acc[pass] += b;
    }
    for (int k = 0; k < 5; k++) per_pass[k] = acc[k];
    return 0;
}

// Exact |B2(v)| per source (brute force with a marker array), the value the
// estimate above is checked against.
int orc_two_hop_exact(const uint64_t* rp, uint64_t nv, const uint32_t* cols, const uint32_t* sources, uint64_t ns,
                      int64_t* out) {
    for (uint64_t i = 0; i < ns; i++)
        if (sources[i] >= nv) return -2;
    std::vector<uint64_t> mark(nv, ~0ull);
    for (uint64_t i = 0; i < ns; i++) {
        const uint32_t v = sources[i];
        int64_t c = 0;
        auto touch = [&](uint32_t w) {
            if (mark[w] != i) {
                mark[w] = i;
                c++;
            }
        };
        touch(v);
        for (uint64_t e = rp[v]; e < rp[v + 1]; e++) {
            const uint32_t u = cols[e];
            touch(u);
            for (uint64_t f = rp[u]; f < rp[u + 1]; f++) touch(cols[f]);
        }
        out[i] = c;
    }
    return 0;
}

}  // extern "C"
