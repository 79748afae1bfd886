// host/graph.cpp -- CSR graph files, preprocessing, degeneracy recode and
// view-bin maps (the data-side host code the device path is fed from).
//
// Reference behaviour restated here:
//   .bin layout           common/graph.h:360-398 (read), :539-562 (write)
//   Preprocess            common/graph.h:720-811
//   kcore recode          kmeans/kcore.cpp:39-191 (peel, shell degree sort, recode)
//   kcore bin map         kmeans/kcore.cpp:193-229
//   .vbmap text           view/view_bin_manager.h:42-46
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <omp.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>

#include "common.hpp"

namespace g2h {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

static bool g_progress = false;
void progress(const char* fmt, ...) {
    if (!g_progress) return;
    static const auto t0 = std::chrono::steady_clock::now();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "[g2h %8.1fs] ", s);
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fputc('\n', stderr);
    std::fflush(stderr);
}

void csr_from_edges(uint64_t nv, const std::vector<uint32_t>& a, const std::vector<uint32_t>& b,
                    g2h_graph& g, int threads) {
    const uint64_t m = a.size();
    std::vector<uint64_t> deg(nv + 1, 0);
    for (uint64_t i = 0; i < m; i++) {
        if (a[i] == b[i]) continue;
        deg[a[i]]++;
        deg[b[i]]++;
    }
    std::vector<uint64_t> off(nv + 1, 0);
    for (uint64_t v = 0; v < nv; v++) off[v + 1] = off[v] + deg[v];
    progress("csr: %llu pairs counted", (unsigned long long)m);
    std::vector<uint32_t> tmp(off[nv]);
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
    for (uint64_t i = 0; i < m; i++) {
        if (a[i] == b[i]) continue;
        tmp[pos[a[i]]++] = b[i];
        tmp[pos[b[i]]++] = a[i];
    }
    progress("csr: %llu entries scattered", (unsigned long long)off[nv]);
    std::vector<uint64_t> uniq(nv, 0);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads)
    for (int64_t v = 0; v < (int64_t)nv; v++) {
        auto s = tmp.begin() + off[v], e = tmp.begin() + off[v + 1];
        std::sort(s, e);
        uniq[v] = std::unique(s, e) - s;
    }
    g.nv = nv;
    g.rp.assign(nv + 1, 0);
    for (uint64_t v = 0; v < nv; v++) g.rp[v + 1] = g.rp[v] + uniq[v];
    g.ne = g.rp[nv];
    g.cols.resize(g.ne);
#pragma omp parallel for schedule(dynamic, 1024) num_threads(threads)
    for (int64_t v = 0; v < (int64_t)nv; v++)
        std::copy(tmp.begin() + off[v], tmp.begin() + off[v] + uniq[v], g.cols.begin() + g.rp[v]);
}

}  // namespace g2h

using namespace g2h;

extern "C" {

const char* g2h_last_error(void) { return g_err.c_str(); }

void g2h_set_progress(int on) { g_progress = on != 0; }

int g2h_graph_read_bin(const char* path, g2h_graph** out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { set_error(std::string("cannot open ") + path); return -1; }
    uint64_t hdr[4];
    if (std::fread(hdr, 8, 4, f) != 4) { std::fclose(f); set_error("short header"); return -2; }
    if (hdr[0] != 4 || hdr[1] != 8) { std::fclose(f); set_error("uintV/uintE size mismatch"); return -3; }
    // The header's sizes must account for the file exactly before anything is
    // allocated from them, so a damaged header cannot ask for a huge buffer.
    const uint64_t nv = hdr[2], ne = hdr[3];
    long long fsize = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) fsize = std::ftell(f);
    const bool sizes_ok = nv < 0xffffffffull && ne <= (1ull << 60) && fsize >= 0 &&
                          (uint64_t)fsize == 32 + 8 * (nv + 1) + 4 * ne;
    if (!sizes_ok || std::fseek(f, 32, SEEK_SET) != 0) {
        std::fclose(f);
        set_error("malformed .bin: header sizes do not match the file");
        return -4;
    }
    auto* g = new g2h_graph;
    g->nv = nv;
    g->ne = ne;
    g->rp.resize(nv + 1);
    g->cols.resize(ne);
    size_t r1 = std::fread(g->rp.data(), 8, nv + 1, f);
    size_t r2 = std::fread(g->cols.data(), 4, ne, f);
    std::fclose(f);
    bool ok = r1 == nv + 1 && r2 == ne && g->rp[nv] == ne;
    for (uint64_t v = 0; ok && v < nv; v++) ok = g->rp[v] <= g->rp[v + 1];
    for (uint64_t j = 0; ok && j < ne; j++) ok = g->cols[j] < nv;
    if (!ok) {
        delete g;
        set_error("malformed .bin body");
        return -4;
    }
    *out = g;
    return 0;
}

int g2h_graph_write_bin(const g2h_graph* g, const char* path) {
    FILE* f = std::fopen(path, "wb");
    if (!f) { set_error(std::string("cannot create ") + path); return -1; }
    uint64_t hdr[4] = {4, 8, g->nv, g->ne};
    size_t w = std::fwrite(hdr, 8, 4, f);
    w += std::fwrite(g->rp.data(), 8, g->nv + 1, f);
    w += std::fwrite(g->cols.data(), 4, g->ne, f);
    std::fclose(f);
    if (w != 4 + g->nv + 1 + g->ne) { set_error("short write"); return -2; }
    return 0;
}

int g2h_graph_from_csr(const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint64_t ne,
                       g2h_graph** out) {
    if (rp[nv] != ne) { set_error("row_ptrs[nv] != ne"); return -1; }
    auto* g = new g2h_graph;
    g->nv = nv;
    g->ne = ne;
    g->rp.assign(rp, rp + nv + 1);
    g->cols.assign(cols, cols + ne);
    *out = g;
    return 0;
}

int g2h_graph_from_edges(const uint32_t* src, const uint32_t* dst, uint64_t m, uint64_t nv,
                         g2h_graph** out) {
    std::vector<uint32_t> a(src, src + m), b(dst, dst + m);
    for (uint64_t i = 0; i < m; i++)
        if (a[i] >= nv || b[i] >= nv) { set_error("edge endpoint >= nv"); return -1; }
    auto* g = new g2h_graph;
    csr_from_edges(nv, a, b, *g, omp_get_max_threads());
    *out = g;
    return 0;
}

uint64_t g2h_graph_nv(const g2h_graph* g) { return g->nv; }
uint64_t g2h_graph_ne(const g2h_graph* g) { return g->ne; }
const uint64_t* g2h_graph_row_ptrs(const g2h_graph* g) { return g->rp.data(); }
const uint32_t* g2h_graph_cols(const g2h_graph* g) { return g->cols.data(); }
uint64_t g2h_graph_max_degree(const g2h_graph* g) {
    uint64_t m = 0;
    for (uint64_t v = 0; v < g->nv; v++) m = std::max(m, g->rp[v + 1] - g->rp[v]);
    return m;
}
void g2h_graph_free(g2h_graph* g) { delete g; }

int g2h_graph_preprocess(g2h_graph* g) {
    progress("preprocess: %llu vertices, %llu entries", (unsigned long long)g->nv, (unsigned long long)g->ne);
    // pass 1: drop self loops and parallel edges (rows are sorted)
    std::vector<uint64_t> nrp(g->nv + 1, 0);
    std::vector<uint32_t> ncols;
    ncols.reserve(g->ne);
    for (uint64_t u = 0; u < g->nv; u++) {
        nrp[u] = ncols.size();
        for (uint64_t j = g->rp[u]; j < g->rp[u + 1]; j++) {
            const uint32_t v = g->cols[j];
            const bool parallel = j > g->rp[u] && v == g->cols[j - 1];
            if (!parallel && v != u) ncols.push_back(v);
        }
    }
    nrp[g->nv] = ncols.size();
    // pass 2: drop degree-0 vertices, relabel densely preserving order
    std::vector<uint32_t> nid(g->nv);
    uint64_t next = 0;
    std::vector<uint64_t> frp;
    frp.reserve(g->nv + 1);
    for (uint64_t u = 0; u < g->nv; u++) {
        if (nrp[u] == nrp[u + 1]) { nid[u] = (uint32_t)g->nv; continue; }
        nid[u] = (uint32_t)next++;
        frp.push_back(nrp[u]);
    }
    frp.push_back(ncols.size());
    for (auto& c : ncols) c = nid[c];
    g->nv = next;
    g->ne = ncols.size();
    g->rp.swap(frp);
    g->cols.swap(ncols);
    return 0;
}

int g2h_graph_kcore_recode(g2h_graph* g, uint32_t* order_out, uint64_t* shell_offsets_out,
                           uint64_t* num_shells, uint32_t* degeneracy) {
    const uint64_t V = g->nv;
    std::vector<uint32_t> deg(V), buffer(V), order(V);
    for (uint64_t v = 0; v < V; v++) deg[v] = (uint32_t)(g->rp[v + 1] - g->rp[v]);
    std::vector<uint64_t> shells{0};
    uint64_t count = 0;
    uint32_t level = 0;
    // level-synchronous peeling (kcore.cpp:39-87, 111-119)
    while (count < V) {
        uint64_t tail = 0;
        for (uint64_t v = 0; v < V; v++)
            if (deg[v] == level) buffer[tail++] = (uint32_t)v;
        for (uint64_t i = 0; i < tail; i++) {
            const uint32_t v = buffer[i];
            for (uint64_t j = g->rp[v]; j < g->rp[v + 1]; j++) {
                const uint32_t u = g->cols[j];
                if (deg[u] > level) {
                    deg[u]--;
                    if (deg[u] == level) buffer[tail++] = u;
                }
            }
        }
        for (uint64_t i = 0; i < tail; i++) order[count + i] = buffer[i];
        count += tail;
        level++;
        shells.push_back(count);
        if (level % 32 == 0) progress("kcore: level %u, %llu of %llu peeled", level, (unsigned long long)count,
                                      (unsigned long long)V);
    }
    // sort each shell by degree (kcore.cpp:126-151); std::sort on the same
    // sequence with the same comparator reproduces the reference's order
    const uint64_t* rp = g->rp.data();
    auto by_deg = [rp](uint32_t a, uint32_t b) { return (rp[a + 1] - rp[a]) < (rp[b + 1] - rp[b]); };
    for (size_t i = 0; i + 1 < shells.size(); i++)
        std::sort(order.begin() + shells[i], order.begin() + shells[i + 1], by_deg);
    std::vector<uint32_t> newid(V);
    for (uint64_t i = 0; i < V; i++) newid[order[i]] = (uint32_t)i;
    // recode (kcore.cpp:153-191)
    std::vector<uint64_t> nrp(V + 1, 0);
    for (uint64_t v = 0; v < V; v++) nrp[newid[v] + 1] = g->rp[v + 1] - g->rp[v];
    for (uint64_t v = 0; v < V; v++) nrp[v + 1] += nrp[v];
    std::vector<uint32_t> ncols(g->ne);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)V; v++) {
        const uint64_t s = nrp[newid[v]];
        uint64_t k = s;
        for (uint64_t j = g->rp[v]; j < g->rp[v + 1]; j++) ncols[k++] = newid[g->cols[j]];
        std::sort(ncols.begin() + s, ncols.begin() + k);
    }
    g->rp.swap(nrp);
    g->cols.swap(ncols);
    if (order_out) std::memcpy(order_out, newid.data(), V * sizeof(uint32_t));
    if (shell_offsets_out) std::memcpy(shell_offsets_out, shells.data(), shells.size() * sizeof(uint64_t));
    if (num_shells) *num_shells = shells.size() - 1;
    if (degeneracy) *degeneracy = level - 1;
    return 0;
}

int g2h_kcore_bins(const uint64_t* shells, uint64_t num_shells, uint32_t nb, uint32_t* bins, uint64_t nv) {
    if (nb == 0) { set_error("nb must be >= 1"); return -1; }
    uint64_t k = 0;
    for (uint64_t i = 0; i < num_shells; i++) {
        uint64_t window = shells[i + 1] - shells[i];
        const uint64_t share = window / nb + 1;
        for (uint32_t j = 0; j < nb && window > 0; j++)
            for (uint64_t t = 0; t < share && window > 0; t++) {
                if (k >= nv) { set_error("shell offsets exceed nv"); return -2; }
                bins[k++] = j;
                window--;
            }
    }
    if (k != nv) { set_error("shell offsets do not cover nv"); return -3; }
    return 0;
}

int g2h_write_vbmap(const char* path, uint32_t nbins, const uint32_t* bins, uint64_t nv) {
    FILE* f = std::fopen(path, "w");
    if (!f) { set_error(std::string("cannot create ") + path); return -1; }
    std::fprintf(f, "%u ", nbins);
    for (uint64_t v = 0; v < nv; v++) std::fprintf(f, "%u ", bins[v]);
    std::fclose(f);
    return 0;
}

int g2h_read_vbmap(const char* path, uint32_t* nbins, uint32_t* bins, uint64_t nv) {
    std::ifstream in(path);
    if (!in) { set_error(std::string("cannot open ") + path); return -1; }
    if (!(in >> *nbins)) { set_error("vbmap has no bin count"); return -2; }
    for (uint64_t v = 0; v < nv; v++) {
        if (!(in >> bins[v])) { set_error("vbmap shorter than vertex count"); return -2; }
        if (bins[v] >= *nbins) { set_error("vbmap bin id out of range"); return -3; }
    }
    return 0;
}

}  // extern "C"
