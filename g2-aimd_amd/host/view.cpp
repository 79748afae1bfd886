// host/view.cpp -- view-bin materialisation (view/view_bin.h:41-159).
//
// A view of a bin keeps the full rows of every vertex within hop-1 steps of a
// source whose degree is >= root_degree; all other rows are empty.  Row
// offsets stay indexed by global vertex id (V+1 entries), as in the
// reference's ViewBinHolder.  The closure is computed breadth first; the
// reference's depth-first ExtendFromSource (view/view_bin.h:41-58) marks
// vertices visited on first touch and can miss vertices at hop >= 3 when a
// vertex is first reached by a longer path.  For hop <= 2 (every pattern the
// benchmark workloads use) both give the same vertex set.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

extern "C" int g2h_view_materialize(const g2h_graph* g, const uint32_t* sources, uint64_t ns, int hop,
                                    uint64_t root_degree, uint64_t* rp_out, uint32_t* cols_out,
                                    uint64_t cols_cap, uint64_t* ne_out) {
    if (hop < 1) { g2h::set_error("hop must be >= 1"); return -1; }
    const uint64_t V = g->nv;
    std::vector<int> dist(V, -1);
    std::vector<uint32_t> frontier, next;
    for (uint64_t i = 0; i < ns; i++) {
        const uint32_t s = sources[i];
        if (s >= V) { g2h::set_error("source id out of range"); return -2; }
        if (root_degree && g->rp[s + 1] - g->rp[s] < root_degree) continue;
        if (dist[s] < 0) { dist[s] = 0; frontier.push_back(s); }
    }
    for (int h = 1; h < hop; h++) {
        next.clear();
        for (uint32_t v : frontier)
            for (uint64_t j = g->rp[v]; j < g->rp[v + 1]; j++) {
                const uint32_t u = g->cols[j];
                if (dist[u] < 0) { dist[u] = h; next.push_back(u); }
            }
        frontier.swap(next);
    }
    uint64_t e = 0;
    for (uint64_t v = 0; v < V; v++) {
        rp_out[v] = e;
        if (dist[v] >= 0) e += g->rp[v + 1] - g->rp[v];
    }
    rp_out[V] = e;
    *ne_out = e;
    if (e > cols_cap) { g2h::set_error("view larger than the holder capacity (-m)"); return -3; }
    for (uint64_t v = 0; v < V; v++)
        if (dist[v] >= 0)
            std::memcpy(cols_out + rp_out[v], g->cols.data() + g->rp[v],
                        (g->rp[v + 1] - g->rp[v]) * sizeof(uint32_t));
    return 0;
}
