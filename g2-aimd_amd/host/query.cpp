// host/query.cpp -- query patterns, symmetry breaking and the matching plan.
//
// Restates (same tables, same orders):
//   preset patterns           common/graph.h:813-1104
//   query CSR build           common/graph.h:127-156
//   Graph::ReMapVertexId      common/graph.h:171-266
//   Graph::GetConditions      common/graph.h:564-686 (+ SetConditions :689-699)
//   Plan::FindRoot .. GenerateStoreStrategy   system/plan.h:16-372
// The reference obtains the automorphism group from bliss generators plus a
// product closure; here the (<= 8-vertex) group is enumerated directly by
// testing every permutation.  GetConditions only depends on the group as a
// set, so the condition pairs are identical; tests/golden holds the
// reference's own tables for the presets and for random patterns.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <queue>
#include <set>
#include <tuple>
#include <vector>

#include "common.hpp"

namespace {

using Rows = std::vector<std::vector<uint32_t>>;

// A pattern is 1..G2_MAXQ vertices whose adjacency masks are symmetric, with
// no self loops and no bits at or above n.
bool valid_pattern(uint32_t n, const uint32_t* qadj) {
    if (n < 1 || n > G2_MAXQ) return false;
    for (uint32_t u = 0; u < n; u++) {
        if ((qadj[u] >> u & 1) || (qadj[u] >> n) != 0) return false;
        for (uint32_t v = 0; v < n; v++)
            if ((qadj[u] >> v & 1) != (qadj[v] >> u & 1)) return false;
    }
    return true;
}

Rows rows_of(uint32_t n, const uint32_t* qadj) {
    Rows r(n);
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t v = 0; v < n; v++)
            if (qadj[u] >> v & 1) r[u].push_back(v);
    return r;
}

bool preset_conn(int t, Rows& c) {
    c.clear();
    auto clique = [&](uint32_t n) {
        c.assign(n, {});
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = 0; j < n; j++) if (i != j) c[i].push_back(j);
    };
    switch (t) {
        case 0: c = {{1, 2}, {0, 2}, {0, 1}}; break;                                   // triangle
        case 1: c = {{1, 3}, {0, 2}, {1, 3}, {0, 2}}; break;                           // square
        case 2: c = {{1, 2, 3}, {0, 2}, {0, 1, 3}, {0, 2}}; break;                     // chordal square
        case 3: c = {{1, 2}, {0, 2}, {0, 1, 3}, {2, 4}, {3}}; break;                   // 2 tails triangle
        case 4: c = {{1, 2}, {0, 2, 3}, {0, 1, 4}, {1, 4}, {2, 3}}; break;             // house
        case 5: c = {{1, 2, 3, 4}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2, 4}, {0, 3}}; break; // chordal house
        case 6: c = {{1, 2, 3}, {0, 2, 3}, {0, 1, 4}, {0, 1, 4}, {2, 3}}; break;       // chordal roof
        case 7: c = {{1, 2, 3, 4}, {0, 3, 2}, {0, 1, 4}, {0, 1}, {0, 2}}; break;       // three triangles
        case 8: c = {{1, 2, 3, 4}, {0, 2, 4}, {0, 1, 3}, {0, 2, 4}, {0, 1, 3}}; break; // solar square
        case 9: c = {{1, 2, 3, 4}, {0, 2, 3, 4}, {0, 1, 3}, {0, 1, 2, 4}, {0, 1, 3}}; break; // near 5 clique
        case 10: c = {{1, 2, 3, 4, 5}, {0, 2}, {0, 1, 3}, {0, 2, 4}, {0, 3, 5}, {0, 4}}; break; // four triangles
        case 11: c = {{1, 2, 3, 5}, {0, 2, 3, 4}, {0, 1, 4, 5}, {0, 1}, {1, 2}, {0, 2}}; break; // one in three
        case 12: c = {{1, 2}, {0, 2, 3, 4, 5}, {0, 1, 3, 4, 5}, {1, 2, 4}, {1, 2, 3, 5}, {1, 2, 4}}; break;
        case 13: c = {{1, 2}, {0, 3}, {0, 3, 4, 5}, {1, 2, 4, 5}, {2, 3}, {2, 3}}; break; // square on top
        case 14: c = {{1, 2, 3, 4, 5}, {0, 2, 3, 5}, {0, 1, 3, 5}, {0, 1, 2, 4, 5, 6}, {0, 3, 5},
                      {0, 1, 2, 3, 4, 6}, {3, 5}}; break;                                  // near 7 clique
        case 15: c = {{1, 2, 3, 4}, {0, 2, 3, 4}, {0, 1, 3, 4}, {0, 1, 2, 4, 5, 6}, {0, 1, 2, 3, 5, 6},
                      {3, 4}, {3, 4}}; break;                                              // 5 clique on top
        case 16: c = {{1, 2}, {0, 3}, {0, 4}, {1, 4}, {2, 3}}; break;                     // 5 circles
        case 17: c = {{1, 2}, {0, 3}, {0, 4}, {1, 5}, {2, 5}, {3, 4}}; break;             // 6 circles
        case 18: c = {{1, 2, 4}, {0, 2, 5}, {0, 1, 3}, {2, 4, 5}, {0, 3, 5}, {1, 3, 4}}; break; // hourglass
        case 23: clique(3); break;
        case 24: clique(4); break;
        case 25: clique(5); break;
        case 26: clique(6); break;
        case 27: clique(7); break;
        default: return false;
    }
    return true;
}

// Graph::ReMapVertexId: BFS from the first maximum-degree vertex; each BFS
// level is ordered by (degree desc, visited-neighbour count desc, id asc)
Rows remap(const Rows& r) {
    const uint32_t n = r.size();
    uint32_t root = 0;
    size_t maxd = 0;
    for (uint32_t i = 0; i < n; i++)
        if (r[i].size() > maxd) { maxd = r[i].size(); root = i; }
    std::vector<bool> vis(n, false);
    std::queue<uint32_t> q;
    q.push(root);
    vis[root] = true;
    uint32_t nid = 0;
    std::vector<uint32_t> o2n(n, 0), n2o(n, 0);
    while (!q.empty()) {
        const size_t size = q.size();
        std::vector<uint32_t> level;
        for (size_t i = 0; i < size; i++) {
            uint32_t f = q.front();
            level.push_back(f);
            q.pop();
            for (uint32_t ne : r[f])
                if (!vis[ne]) { vis[ne] = true; q.push(ne); }
        }
        std::vector<std::tuple<size_t, size_t, uint32_t>> w;
        for (uint32_t v : level) {
            size_t conn = 0;
            for (uint32_t ne : r[v]) conn += vis[ne];
            w.emplace_back(r[v].size(), conn, v);
        }
        std::sort(w.begin(), w.end(), [](const auto& a, const auto& b) {
            if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) > std::get<0>(b);
            if (std::get<1>(a) != std::get<1>(b)) return std::get<1>(a) > std::get<1>(b);
            return std::get<2>(a) < std::get<2>(b);
        });
        for (auto& t : w) { o2n[std::get<2>(t)] = nid; n2o[nid] = std::get<2>(t); nid++; }
    }
    Rows out(n);
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t ne : r[i]) out[o2n[i]].push_back(o2n[ne]);
    }
    for (auto& row : out) std::sort(row.begin(), row.end());
    return out;
}

bool adjacent(const Rows& r, uint32_t a, uint32_t b) {
    return std::binary_search(r[a].begin(), r[a].end(), b);
}

std::vector<std::vector<uint32_t>> automorphisms(const Rows& r) {
    const uint32_t n = r.size();
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::vector<std::vector<uint32_t>> out;
    do {
        bool ok = true;
        for (uint32_t a = 0; a < n && ok; a++) {
            if (r[a].size() != r[perm[a]].size()) ok = false;
            for (uint32_t b : r[a]) if (ok && !adjacent(r, perm[a], perm[b])) ok = false;
        }
        if (ok) out.push_back(perm);
    } while (std::next_permutation(perm.begin(), perm.end()));
    return out;
}

std::map<uint32_t, std::set<uint32_t>> eq_classes(uint32_t n, const std::vector<std::vector<uint32_t>>& aut) {
    std::map<uint32_t, std::set<uint32_t>> ec;
    for (uint32_t i = 0; i < n; i++) {
        std::set<uint32_t> e;
        for (auto& p : aut) e.insert(p[i]);
        if (e.empty()) e.insert(i);
        const uint32_t rep = *e.begin();
        ec[rep].insert(e.begin(), e.end());
    }
    return ec;
}

std::vector<std::pair<uint32_t, uint32_t>> conditions(const Rows& r) {
    const uint32_t n = r.size();
    auto aut = automorphisms(r);
    auto ec = eq_classes(n, aut);
    std::vector<std::pair<uint32_t, uint32_t>> res;
    auto big = [](const auto& e) { return e.second.size() > 1; };
    auto it = std::find_if(ec.cbegin(), ec.cend(), big);
    while (it != ec.cend()) {
        const auto& e = it->second;
        const uint32_t n0 = *e.cbegin();
        for (auto& p : aut) {
            uint32_t best = 0;
            bool first = true;
            for (auto x = std::next(e.cbegin()); x != e.cend(); ++x)
                if (first || p[*x] < p[best]) { best = *x; first = false; }
            res.emplace_back(n0, best);
        }
        aut.erase(std::remove_if(aut.begin(), aut.end(), [n0](const auto& p) { return p[n0] != n0; }), aut.end());
        ec = eq_classes(n, aut);
        it = std::find_if(ec.cbegin(), ec.cend(), big);
    }
    std::sort(res.begin(), res.end());
    res.erase(std::unique(res.begin(), res.end()), res.end());
    return res;
}

// Plan::GetMinHopWithBFS (system/plan.h:16-49)
int min_hop(const Rows& r, uint32_t root) {
    const uint32_t n = r.size();
    std::vector<bool> vis(n, false);
    std::vector<int> lvl(n, -1);
    int maxv = -1, maxe = -1;
    std::queue<uint32_t> q;
    q.push(root);
    vis[root] = true;
    lvl[root] = 0;
    while (!q.empty()) {
        uint32_t f = q.front();
        q.pop();
        maxv = lvl[f];
        for (uint32_t j : r[f]) {
            if (!vis[j]) { vis[j] = true; lvl[j] = lvl[f] + 1; q.push(j); }
            else if (lvl[j] == lvl[f]) maxe = lvl[j];
        }
    }
    return maxv + (maxv == maxe ? 1 : 0);
}

}  // namespace

extern "C" {

int g2h_query_preset(int preset, uint32_t* n, uint32_t qadj[G2_MAXQ]) {
    Rows c;
    if (!preset_conn(preset, c)) { g2h::set_error("unknown preset pattern"); return -1; }
    // Graph(QUERY) builds rows from the conn lists and sorts them
    Rows r(c.size());
    for (uint32_t i = 0; i < c.size(); i++)
        for (uint32_t j : c[i]) r[j].push_back(i);
    for (auto& row : r) std::sort(row.begin(), row.end());
    Rows m = remap(r);
    *n = m.size();
    for (uint32_t u = 0; u < G2_MAXQ; u++) qadj[u] = 0;
    for (uint32_t u = 0; u < m.size(); u++)
        for (uint32_t v : m[u]) qadj[u] |= 1u << v;
    return 0;
}

int g2h_query_remap(uint32_t n, const uint32_t qadj_in[G2_MAXQ], uint32_t qadj_out[G2_MAXQ]) {
    if (n < 1 || n > G2_MAXQ) { g2h::set_error("pattern size out of range"); return -1; }
    if (!valid_pattern(n, qadj_in)) { g2h::set_error("pattern adjacency must be symmetric, without self loops"); return -1; }
    Rows m = remap(rows_of(n, qadj_in));
    for (uint32_t u = 0; u < G2_MAXQ; u++) qadj_out[u] = 0;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t v : m[u]) qadj_out[u] |= 1u << v;
    return 0;
}

int g2h_query_conditions(uint32_t n, const uint32_t qadj[G2_MAXQ], uint32_t* npairs,
                         uint32_t pairs[2 * G2_MAXQ * G2_MAXQ]) {
    if (n < 1 || n > G2_MAXQ) { g2h::set_error("pattern size out of range"); return -1; }
    if (!valid_pattern(n, qadj)) { g2h::set_error("pattern adjacency must be symmetric, without self loops"); return -1; }
    auto c = conditions(rows_of(n, qadj));
    *npairs = c.size();
    for (size_t k = 0; k < c.size(); k++) { pairs[2 * k] = c[k].first; pairs[2 * k + 1] = c[k].second; }
    return 0;
}

uint64_t g2h_query_automorphisms(uint32_t n, const uint32_t qadj[G2_MAXQ]) {
    if (!valid_pattern(n, qadj)) return 0;
    return automorphisms(rows_of(n, qadj)).size();
}

// Isomorphism-invariant code: the smallest upper-triangle adjacency bit
// string over all relabellings (n <= 8, so at most 8! = 40320 of them), with
// n in the top byte.  Equal codes <=> isomorphic patterns.
uint64_t g2h_query_canonical(uint32_t n, const uint32_t qadj[G2_MAXQ]) {
    if (!valid_pattern(n, qadj)) return 0;
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    uint64_t best = ~0ull;
    do {
        uint64_t code = 0;
        uint32_t bit = 0;
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = i + 1; j < n; j++, bit++)
                if (qadj[perm[i]] >> perm[j] & 1) code |= 1ull << bit;
        best = std::min(best, code);
    } while (std::next_permutation(perm.begin(), perm.end()));
    return ((uint64_t)n << 56) | best;
}

int g2h_plan_build(uint32_t n, const uint32_t qadj[G2_MAXQ], g2_plan* P) {
    if (n < 2 || n > G2_MAXQ) { g2h::set_error("pattern size out of range"); return -1; }
    if (!valid_pattern(n, qadj)) { g2h::set_error("pattern adjacency must be symmetric, without self loops"); return -1; }
    const Rows r = rows_of(n, qadj);
    for (uint32_t u = 0; u < n; u++)
        if (r[u].empty()) { g2h::set_error("pattern is not connected"); return -2; }
    *P = g2_plan{};
    P->n = n;
    // SetConditions (common/graph.h:689-699): per vertex, in pair order
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> order(n);
    for (auto& c : conditions(r)) {
        order[c.first].push_back({0u, c.second});   // LESS_THAN
        order[c.second].push_back({1u, c.first});   // LARGER_THAN
    }
    // FindRoot (plan.h:51-73)
    uint32_t root = 0;
    int hop = 10;
    size_t rdeg = 0;
    for (uint32_t u = 0; u < n; u++) {
        const int t = min_hop(r, u);
        if (t < hop) { root = u; hop = t; rdeg = r[u].size(); }
        else if (t == hop && rdeg < r[u].size()) { root = u; hop = t; rdeg = r[u].size(); }
    }
    // reachability check (a disconnected pattern never reaches every vertex)
    {
        std::vector<bool> vis(n, false);
        std::queue<uint32_t> q;
        q.push(root);
        vis[root] = true;
        uint32_t seen = 1;
        while (!q.empty()) {
            uint32_t f = q.front();
            q.pop();
            for (uint32_t j : r[f]) if (!vis[j]) { vis[j] = true; seen++; q.push(j); }
        }
        if (seen != n) { g2h::set_error("pattern is not connected"); return -2; }
    }
    P->hop = hop;
    P->root_degree = rdeg;
    // GenerateSearchSequence (plan.h:75-123)
    {
        std::vector<std::tuple<int, int, uint32_t>> w(n);
        std::vector<bool> vis(n, false);
        vis[root] = true;
        std::queue<uint32_t> q;
        q.push(root);
        int h = 0;
        w[root] = {h, (int)r[root].size(), root};
        while (!q.empty()) {
            h++;
            const size_t size = q.size();
            for (size_t i = 0; i < size; i++) {
                uint32_t f = q.front();
                q.pop();
                for (uint32_t j : r[f])
                    if (!vis[j]) { vis[j] = true; q.push(j); w[j] = {h, (int)r[j].size(), j}; }
            }
        }
        std::sort(w.begin(), w.end(), [](const auto& a, const auto& b) {
            if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) < std::get<0>(b);
            if (std::get<1>(a) != std::get<1>(b)) return std::get<1>(a) > std::get<1>(b);
            return std::get<2>(a) < std::get<2>(b);
        });
        for (uint32_t i = 0; i < n; i++) { P->seq[i] = std::get<2>(w[i]); P->rev[P->seq[i]] = i; }
    }
    // GenerateBackwardNeighbor (plan.h:125-181)
    {
        std::vector<bool> vis(n, false);
        vis[P->seq[0]] = true;
        for (uint32_t i = 1; i < n; i++) {
            const uint32_t v = P->seq[i];
            for (uint32_t nv : r[v])
                if (vis[nv]) P->bn[i][P->bn_count[i]++] = nv;
            vis[v] = true;
        }
        P->shared[0] = 0;
        for (uint32_t i = 1; i < n; i++) {
            P->shared[i] = 0;
            if (P->bn_count[i] > 2) P->shared[i] = 1;
            else if (P->bn_count[i] == 2) {
                const uint32_t last = P->seq[i - 1];
                if (P->bn[i][0] != last && P->bn[i][1] != last) P->shared[i] = 1;
            }
        }
    }
    // GeneratePreAfterBackwardNeighbor (plan.h:183-214)
    for (uint32_t i = 2; i < n; i++) {
        if (!P->shared[i]) continue;
        const uint32_t v = P->seq[i];
        for (uint32_t nv : r[v]) {
            if (P->rev[nv] + 1 < P->rev[v]) P->pre_bn[i][P->pre_bn_count[i]++] = nv;
            else if (P->rev[nv] + 1 == P->rev[v]) P->after_bn[i][P->after_bn_count[i]++] = nv;
        }
    }
    // GenerateUsefulOrder (plan.h:216-325); rev[i]-1 is unsigned in the
    // reference, so for the root (rev 0) every j qualifies as "pre"
    auto add = [&](uint32_t i, uint32_t j, uint32_t* cnt, uint32_t (*tab)[G2_MAXQ][2]) {
        uint32_t op = 2;  // NON_EQUAL unless an order condition exists
        for (auto& c : order[i]) if (c.second == j) { op = c.first; break; }
        tab[i][cnt[i]][0] = op;
        tab[i][cnt[i]][1] = j;
        cnt[i]++;
    };
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < n; j++)
            if (P->rev[i] > P->rev[j]) add(i, j, P->cond_count, P->cond);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < n; j++)
            if ((uint32_t)(P->rev[i] - 1u) > P->rev[j]) add(i, j, P->pre_cond_count, P->pre_cond);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < n; j++)
            if ((uint32_t)(P->rev[i] - 1u) == P->rev[j]) add(i, j, P->after_cond_count, P->after_cond);
    // GenerateStoreStrategy (plan.h:327-372)
    for (uint32_t i = 0; i < n; i++) P->strategy[i] = P->shared[i] ? 1 : 0;
    P->strategy[n] = 2;
    for (uint32_t i = 0; i < n; i++) P->moving_lvl[i] = 1;
    for (uint32_t i = 1; i < n; i++) {
        uint32_t same = 1;
        for (uint32_t j = i + 1; j < n; j++) {
            bool sb = P->bn_count[i] == P->bn_count[j];
            for (uint32_t k = 0; sb && k < P->bn_count[i]; k++)
                if (P->bn[i][k] != P->bn[j][k]) sb = false;
            if (!sb) break;
            same++;
        }
        P->moving_lvl[i] = same;
    }
    return 0;
}

int g2h_query_random(uint32_t n, double density, uint64_t seed, uint32_t qadj[G2_MAXQ]) {
    if (n < 2 || n > G2_MAXQ) { g2h::set_error("pattern size out of range"); return -1; }
    g2h::Rng r(seed);
    for (uint32_t u = 0; u < G2_MAXQ; u++) qadj[u] = 0;
    // random spanning tree: vertex i attaches to a uniform earlier vertex of a
    // random permutation
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    for (uint32_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[r.below(i + 1)]);
    uint32_t m = 0;
    for (uint32_t i = 1; i < n; i++) {
        const uint32_t a = perm[i], b = perm[r.below(i)];
        qadj[a] |= 1u << b;
        qadj[b] |= 1u << a;
        m++;
    }
    const uint32_t full = n * (n - 1) / 2;
    uint32_t target = (uint32_t)std::llround(density * full);
    target = std::max(target, n - 1);
    target = std::min(target, full);
    std::vector<std::pair<uint32_t, uint32_t>> free;
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t b = a + 1; b < n; b++)
            if (!(qadj[a] >> b & 1)) free.push_back({a, b});
    for (size_t i = free.size(); i > 1; i--) std::swap(free[i - 1], free[r.below(i)]);
    for (size_t k = 0; m < target && k < free.size(); k++, m++) {
        qadj[free[k].first] |= 1u << free[k].second;
        qadj[free[k].second] |= 1u << free[k].first;
    }
    return 0;
}

}  // extern "C"
