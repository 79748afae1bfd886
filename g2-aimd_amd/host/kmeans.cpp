// host/kmeans.cpp -- k-means view binning and view packing, the reference's
// offline input tools kmeans/kmeans.cpp:53-795 and kmeans/view_pack.cpp:123-365
// (with kmeans/kmeans_util.h and kmeans/view_bin.h), restated.
//
// k-means clusters vertices by the cosine similarity of their adjacency
// rows, each row scaled by 1/sqrt(degree): a centroid is the dense mean of
// its members' scaled rows over all V coordinates, and
//   sim(c, v) = sum_{u in N(v)} c[u] / sqrt(deg v) / |c|.
// Differences by design: the initialisation is seeded (the reference seeds
// std::rand with time()), ties break deterministically, clusters update with
// a fixed summation order, so the same seed gives the same bins.
//
// View packing puts each cluster's vertices, in id order (or by decreasing
// degree), into open view bins: a vertex goes to the open bin whose hop
// closure grows least while staying within max_entries row entries; when
// none fits, a new bin opens, closing the largest open one if pool_size are
// open.  Bins are then merged next-fit in increasing size.  A vertex whose
// closure alone exceeds the limit gets a bin of its own (the reference
// drops it), and a tentative insertion that is rolled back clears the marks
// it set (the reference's recall clears `added` twice and never `visited`).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include <omp.h>

#include "common.hpp"

namespace {

struct Csr {
    const uint64_t* rp;
    const uint32_t* cols;
    uint64_t nv;
    uint64_t deg(uint32_t v) const { return rp[v + 1] - rp[v]; }
};

struct Centroid {
    std::vector<double> c;
    double norm = 0;  // sqrt of the sum of squares
};

inline double row_scale(const Csr& g, uint32_t v) {
    const uint64_t d = g.deg(v);
    return d ? 1.0 / std::sqrt((double)d) : 0.0;
}

inline double similarity(const Csr& g, const Centroid& c, uint32_t v) {
    if (c.norm == 0 || g.deg(v) == 0) return 0;
    double s = 0;
    for (uint64_t e = g.rp[v]; e < g.rp[v + 1]; e++) s += c.c[g.cols[e]];
    return s * row_scale(g, v) / c.norm;
}

void finish(Centroid& c, uint64_t members) {
    double sq = 0;
    for (double& x : c.c) {
        if (members) x /= (double)members;
        sq += x * x;
    }
    c.norm = std::sqrt(sq);
}

void centroid_of(const Csr& g, const std::vector<uint32_t>& members, Centroid& c) {
    std::fill(c.c.begin(), c.c.end(), 0.0);
    for (uint32_t v : members) {
        const double w = row_scale(g, v);
        for (uint64_t e = g.rp[v]; e < g.rp[v + 1]; e++) c.c[g.cols[e]] += w;
    }
    finish(c, members.size());
}

// k-means over the vertices `vids` (a subset of the graph); assign[i] for vids[i]
void kmeans_subset(const Csr& g, const std::vector<uint32_t>& vids, uint32_t K, uint32_t iters, uint64_t seed,
                   std::vector<uint32_t>& assign) {
    const uint64_t n = vids.size();
    assign.assign(n, 0);
    if (n == 0 || K <= 1) return;
    std::vector<Centroid> cen(K);
    for (auto& c : cen) c.c.assign(g.nv, 0.0);
    // farthest-first initialisation: the first centroid is a seeded random
    // member; each next one is the member least similar to its most similar
    // chosen centroid (ties: larger degree, then lower position)
    g2h::Rng rng(seed);
    std::vector<double> maxsim(n, -1e300);
    std::vector<char> chosen(n, 0);
    uint64_t pick = rng.below(n);
    for (uint32_t k = 0; k < K; k++) {
        chosen[pick] = 1;
        centroid_of(g, {vids[pick]}, cen[k]);
        if (k + 1 == K) break;
        int nt = omp_get_max_threads();
        std::vector<std::pair<double, uint64_t>> best(nt, {1e300, n});
#pragma omp parallel num_threads(nt)
        {
            auto& b = best[omp_get_thread_num()];
#pragma omp for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; i++) {
                if (chosen[i]) continue;
                maxsim[i] = std::max(maxsim[i], similarity(g, cen[k], vids[i]));
                const bool better = maxsim[i] < b.first ||
                                    (maxsim[i] == b.first && b.second < n &&
                                     (g.deg(vids[i]) > g.deg(vids[b.second]) ||
                                      (g.deg(vids[i]) == g.deg(vids[b.second]) && (uint64_t)i < b.second)));
                if (better) b = {maxsim[i], (uint64_t)i};
            }
        }
        std::pair<double, uint64_t> w{1e300, n};
        for (auto& b : best) {
            if (b.second == n) continue;
            const bool better = w.second == n || b.first < w.first ||
                                (b.first == w.first && (g.deg(vids[b.second]) > g.deg(vids[w.second]) ||
                                                        (g.deg(vids[b.second]) == g.deg(vids[w.second]) &&
                                                         b.second < w.second)));
            if (better) w = b;
        }
        if (w.second == n) {  // fewer members than clusters: the rest stay empty
            for (uint32_t j = k + 1; j < K; j++) cen[j].norm = 0;
            break;
        }
        pick = w.second;
    }
    std::vector<std::vector<uint32_t>> members(K);
    for (uint32_t it = 0; it <= iters; it++) {
        // assignment: the most similar centroid (the lowest index on ties)
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            uint32_t a = 0;
            double best = similarity(g, cen[0], vids[i]);
            for (uint32_t k = 1; k < K; k++) {
                const double s = similarity(g, cen[k], vids[i]);
                if (s > best) {
                    best = s;
                    a = k;
                }
            }
            assign[i] = a;
        }
        if (it == iters) break;
        for (auto& m : members) m.clear();
        for (uint64_t i = 0; i < n; i++) members[assign[i]].push_back(vids[i]);
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t k = 0; k < (int64_t)K; k++) centroid_of(g, members[k], cen[k]);
    }
}

}  // namespace

extern "C" {

int g2h_graph_read_snap(const char* path, g2h_graph** out) {
    if (!path || !out) { g2h::set_error("null argument"); return -1; }
    *out = nullptr;
    FILE* f = std::fopen(path, "r");
    if (!f) { g2h::set_error(std::string("cannot read ") + path); return -2; }
    // every line starting with a digit holds an edge "a b"; ids are shifted
    // by the smallest one; both directions are stored, rows sorted
    // (common/graph.h:268-358); parallel edges and self loops stay for Preprocess
    std::vector<uint32_t> a, b;
    char line[4096];
    unsigned long long lo = ~0ull, hi = 0;
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] < '0' || line[0] > '9') continue;
        unsigned long long x = 0, y = 0;
        if (std::sscanf(line, "%llu %llu", &x, &y) != 2) {
            std::fclose(f);
            g2h::set_error(std::string("malformed edge line in ") + path);
            return -3;
        }
        if (x > 0xfffffffeull || y > 0xfffffffeull) {
            std::fclose(f);
            g2h::set_error("vertex id does not fit in u32");
            return -3;
        }
        a.push_back((uint32_t)x);
        b.push_back((uint32_t)y);
        lo = std::min(lo, std::min(x, y));
        hi = std::max(hi, std::max(x, y));
    }
    std::fclose(f);
    auto* g = new g2h_graph;
    if (a.empty()) {
        g->rp.assign(1, 0);
        *out = g;
        return 0;
    }
    g->nv = hi - lo + 1;
    g->ne = 2 * a.size();
    g->rp.assign(g->nv + 1, 0);
    for (size_t i = 0; i < a.size(); i++) {
        a[i] -= (uint32_t)lo;
        b[i] -= (uint32_t)lo;
        g->rp[a[i] + 1]++;
        g->rp[b[i] + 1]++;
    }
    for (uint64_t v = 0; v < g->nv; v++) g->rp[v + 1] += g->rp[v];
    std::vector<uint64_t> pos(g->rp.begin(), g->rp.end() - 1);
    g->cols.resize(g->ne);
    for (size_t i = 0; i < a.size(); i++) {
        g->cols[pos[a[i]]++] = b[i];
        g->cols[pos[b[i]]++] = a[i];
    }
    for (uint64_t v = 0; v < g->nv; v++) std::sort(g->cols.begin() + g->rp[v], g->cols.begin() + g->rp[v + 1]);
    *out = g;
    return 0;
}

int g2h_kmeans(const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint32_t k1, uint32_t k2, uint32_t iters,
               uint64_t seed, uint32_t* cluster, uint32_t* nclusters) {
    if (!rp || !cluster || !nclusters || (rp[nv] && !cols)) { g2h::set_error("null argument"); return -1; }
    if (k1 == 0 || k2 == 0 || (uint64_t)k1 * k2 > 65536) { g2h::set_error("cluster counts must be 1..65536"); return -1; }
    const Csr g{rp, cols, nv};
    std::vector<uint32_t> all(nv), assign;
    std::iota(all.begin(), all.end(), 0u);
    kmeans_subset(g, all, k1, iters, seed, assign);
    std::vector<uint32_t> global(nv);
    for (uint64_t v = 0; v < nv; v++) global[v] = assign[v] * k2;
    if (k2 > 1) {  // second level: k2 clusters inside each first-level one
        std::vector<std::vector<uint32_t>> part(k1);
        for (uint64_t v = 0; v < nv; v++) part[assign[v]].push_back((uint32_t)v);
        for (uint32_t c = 0; c < k1; c++) {
            std::vector<uint32_t> sub;
            kmeans_subset(g, part[c], k2, iters, seed + 1 + c, sub);
            for (size_t i = 0; i < part[c].size(); i++) global[part[c][i]] = c * k2 + sub[i];
        }
    }
    // clusters numbered by first appearance order of their ids, empty ones removed
    std::vector<uint32_t> remap((uint64_t)k1 * k2, ~0u);
    std::vector<uint64_t> size((uint64_t)k1 * k2, 0);
    for (uint64_t v = 0; v < nv; v++) size[global[v]]++;
    uint32_t n = 0;
    for (uint64_t c = 0; c < size.size(); c++)
        if (size[c]) remap[c] = n++;
    for (uint64_t v = 0; v < nv; v++) cluster[v] = remap[global[v]];
    *nclusters = n;
    return 0;
}

namespace {

// one view bin of the packing: the vertices whose rows it holds (`added`),
// the vertices already expanded (`visited`), its sources and size
struct PackBin {
    std::vector<char> added, visited;
    std::vector<uint32_t> sources;
    uint64_t size = 0;
    std::vector<uint32_t> new_added, new_visited;  // marks of the last tentative insertion
    explicit PackBin(uint64_t nv) : added(nv, 0), visited(nv, 0) {}
    // row entries the hop closure of v adds (depth-first, view/view_bin.h:41-58 order)
    uint64_t grow(const Csr& g, uint32_t v, int hop) {
        uint64_t s = 0;
        if (!added[v]) {
            added[v] = 1;
            new_added.push_back(v);
            s += g.deg(v);
        }
        if (hop > 1) {
            visited[v] = 1;
            new_visited.push_back(v);
            for (uint64_t e = g.rp[v]; e < g.rp[v + 1]; e++)
                if (!visited[g.cols[e]]) s += grow(g, g.cols[e], hop - 1);
        }
        return s;
    }
    void undo() {
        for (uint32_t v : new_added) added[v] = 0;
        for (uint32_t v : new_visited) visited[v] = 0;
    }
};

}  // namespace

int g2h_view_pack(const uint64_t* rp, uint64_t nv, const uint32_t* cols, const uint32_t* cluster, uint32_t nclusters,
                  int hop, uint64_t max_entries, uint32_t pool_size, int by_degree, uint32_t* bins, uint32_t* nbins) {
    if (!rp || !cluster || !bins || !nbins || (rp[nv] && !cols)) { g2h::set_error("null argument"); return -1; }
    if (hop < 1 || pool_size < 1 || nclusters < 1) { g2h::set_error("hop, pool size and cluster count must be >= 1"); return -1; }
    const Csr g{rp, cols, nv};
    std::vector<std::vector<uint32_t>> chunk(nclusters);
    for (uint64_t v = 0; v < nv; v++) {
        if (cluster[v] >= nclusters) { g2h::set_error("cluster id out of range"); return -1; }
        chunk[cluster[v]].push_back((uint32_t)v);
    }
    if (by_degree)
        for (auto& c : chunk) std::stable_sort(c.begin(), c.end(), [&](uint32_t a, uint32_t b) { return g.deg(a) > g.deg(b); });
    // every cluster packs into its own bins (independently, in parallel)
    std::vector<std::vector<PackBin>> packed(nclusters);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < (int64_t)nclusters; c++) {
        auto& out = packed[c];
        std::vector<size_t> open;
        auto create = [&]() {
            if (open.size() >= pool_size) {  // close the largest open bin
                size_t arg = 0;
                for (size_t i = 1; i < open.size(); i++)
                    if (out[open[i]].size > out[open[arg]].size) arg = i;
                std::vector<char>().swap(out[open[arg]].added);  // a closed bin needs no marks
                std::vector<char>().swap(out[open[arg]].visited);
                open.erase(open.begin() + arg);
            }
            out.emplace_back(nv);
            open.push_back(out.size() - 1);
        };
        create();
        for (uint32_t u : chunk[c]) {
            size_t best = SIZE_MAX;
            uint64_t best_inc = 0;
            std::vector<uint64_t> inc(open.size());
            for (size_t i = 0; i < open.size(); i++) {
                PackBin& b = out[open[i]];
                b.new_added.clear();
                b.new_visited.clear();
                inc[i] = b.grow(g, u, hop);
                if (b.size + inc[i] <= max_entries && (best == SIZE_MAX || inc[i] < best_inc)) {
                    best = i;
                    best_inc = inc[i];
                }
            }
            if (best == SIZE_MAX) {
                for (size_t i = 0; i < open.size(); i++) out[open[i]].undo();
                create();
                PackBin& b = out[open.back()];
                b.new_added.clear();
                b.new_visited.clear();
                b.size += b.grow(g, u, hop);  // a fresh bin takes it even beyond the limit
                b.sources.push_back(u);
                continue;
            }
            for (size_t i = 0; i < open.size(); i++)
                if (i != best) out[open[i]].undo();
            PackBin& b = out[open[best]];
            b.size += best_inc;
            b.sources.push_back(u);
        }
        for (auto& b : out) {  // the marks are no longer needed
            std::vector<char>().swap(b.added);
            std::vector<char>().swap(b.visited);
        }
    }
    // merge next-fit in increasing size (ties: creation order)
    std::vector<const PackBin*> all;
    for (auto& p : packed)
        for (auto& b : p) all.push_back(&b);
    std::vector<size_t> order(all.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return all[a]->size < all[b]->size; });
    uint32_t merged = 0;
    uint64_t cur = 0;
    bool first = true;
    for (size_t i : order) {
        const uint64_t sz = all[i]->size;
        if (!first && cur + sz > max_entries) {
            merged++;
            cur = 0;
        }
        first = false;
        cur += sz;
        for (uint32_t v : all[i]->sources) bins[v] = merged;
    }
    *nbins = all.empty() ? 0 : merged + 1;
    return 0;
}

int g2h_write_vbmap_lines(const char* path, uint32_t nbins, const uint32_t* bins, uint64_t nv) {
    if (!path || (nv && !bins)) { g2h::set_error("null argument"); return -1; }
    FILE* f = std::fopen(path, "w");
    if (!f) { g2h::set_error(std::string("cannot write ") + path); return -2; }
    std::fprintf(f, "%u\n", nbins);
    for (uint64_t v = 0; v < nv; v++) std::fprintf(f, "%u\n", bins[v]);
    const bool ok = std::fflush(f) == 0 && !std::ferror(f);
    std::fclose(f);
    if (!ok) { g2h::set_error(std::string("write error on ") + path); return -2; }
    return 0;
}

}  // extern "C"
