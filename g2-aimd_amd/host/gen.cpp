// host/gen.cpp -- seeded synthetic data graphs for the benchmark workloads.
//
// Every generator is deterministic for a given seed and independent of the
// thread count: edge draws are cut into fixed chunks of 2^20 and chunk c draws
// from its own xoshiro256** stream seeded with splitmix64(seed + c).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <omp.h>

#include "common.hpp"

namespace {

using g2h::Rng;

constexpr uint64_t kChunk = 1ull << 20;

inline uint64_t chunk_seed(uint64_t seed, uint64_t c) {
    uint64_t z = seed + 0x9e3779b97f4a7c15ULL * (c + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Vose alias table for sampling an index proportional to w
struct Alias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    explicit Alias(const std::vector<double>& w) {
        const size_t n = w.size();
        prob.resize(n);
        alias.resize(n);
        const double sum = std::accumulate(w.begin(), w.end(), 0.0);
        std::vector<double> p(n);
        std::vector<uint32_t> small, large;
        for (size_t i = 0; i < n; i++) {
            p[i] = w[i] * n / sum;
            (p[i] < 1.0 ? small : large).push_back((uint32_t)i);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t s = small.back(); small.pop_back();
            uint32_t l = large.back(); large.pop_back();
            prob[s] = p[s];
            alias[s] = l;
            p[l] = (p[l] + p[s]) - 1.0;
            (p[l] < 1.0 ? small : large).push_back(l);
        }
        for (uint32_t l : large) { prob[l] = 1.0; alias[l] = l; }
        for (uint32_t s : small) { prob[s] = 1.0; alias[s] = s; }
    }
    uint32_t draw(Rng& r) const {
        const uint64_t i = r.below(prob.size());
        return r.uniform() < prob[i] ? (uint32_t)i : alias[i];
    }
};

// power-law expected-degree sequence, i0 solved by bisection so that the
// unclipped sum equals target_sum
std::vector<double> chung_lu_weights(uint64_t n, double gamma, double w_min, double w_max,
                                     double target_sum) {
    const double ex = -1.0 / (gamma - 1.0);
    auto total = [&](double i0) {
        double s = 0;
        for (uint64_t i = 0; i < n; i++) s += std::max(w_min, w_max * std::pow(1.0 + i / i0, ex));
        return s;
    };
    double lo = 1e-6, hi = (double)n * 1e3;
    for (int it = 0; it < 200; it++) {
        const double mid = std::sqrt(lo * hi);
        if (total(mid) < target_sum) lo = mid; else hi = mid;
        if (hi / lo < 1.0 + 1e-12) break;
    }
    const double i0 = std::sqrt(lo * hi);
    std::vector<double> w(n);
    for (uint64_t i = 0; i < n; i++) w[i] = std::max(w_min, w_max * std::pow(1.0 + i / i0, ex));
    return w;
}

void draw_pairs(const Alias& al, uint64_t m, uint64_t seed, std::vector<uint32_t>& a,
                std::vector<uint32_t>& b, uint64_t base) {
    const uint64_t nch = (m + kChunk - 1) / kChunk;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < (int64_t)nch; c++) {
        Rng r(chunk_seed(seed, c));
        const uint64_t s = c * kChunk, e = std::min(m, s + kChunk);
        for (uint64_t i = s; i < e; i++) {
            a[base + i] = al.draw(r);
            b[base + i] = al.draw(r);
        }
    }
}

// expected number of maximal cliques of G(s,p):
//   sum_k C(s,k) p^{k(k-1)/2} (1-p^k)^{s-k}
double expected_max_cliques(uint32_t s, double p) {
    double tot = 0, logc = 0;
    const double lp = std::log(p);
    for (uint32_t k = 1; k <= s; k++) {
        logc += std::log((double)(s - k + 1)) - std::log((double)k);
        const double t = logc + lp * k * (k - 1) / 2.0 + (s - k) * std::log1p(-std::pow(p, k));
        tot += std::exp(t);
    }
    return tot;
}

}  // namespace

extern "C" {

int g2h_gen_chung_lu(uint64_t n, double gamma, double w_min, double w_max, uint64_t target_edges,
                     uint64_t seed, g2h_graph** out) {
    if (n < 2 || n > 0xffffffffull || gamma <= 1.0) { g2h::set_error("bad Chung-Lu parameters"); return -1; }
    auto w = chung_lu_weights(n, gamma, w_min, w_max, 2.0 * target_edges);
    const double sum = std::accumulate(w.begin(), w.end(), 0.0);
    const uint64_t m = (uint64_t)std::llround(sum / 2.0);
    Alias al(w);
    std::vector<uint32_t> a(m), b(m);
    draw_pairs(al, m, seed, a, b, 0);
    auto* g = new g2h_graph;
    g2h::csr_from_edges(n, a, b, *g, omp_get_max_threads());
    *out = g;
    return 0;
}

int g2h_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                 uint64_t seed, g2h_graph** out) {
    if (scale < 1 || scale > 31) { g2h::set_error("bad R-MAT scale"); return -1; }
    const uint64_t n = 1ull << scale;
    const uint64_t m = (uint64_t)edge_factor * n;
    const double ab = a + b, abc = a + b + c;
    std::vector<uint32_t> src(m), dst(m);
    const uint64_t nch = (m + kChunk - 1) / kChunk;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < (int64_t)nch; ch++) {
        Rng r(chunk_seed(seed, ch));
        const uint64_t s = ch * kChunk, e = std::min(m, s + kChunk);
        for (uint64_t i = s; i < e; i++) {
            uint32_t u = 0, v = 0;
            for (uint32_t l = 0; l < scale; l++) {
                const double x = r.uniform();
                const uint32_t bu = x >= ab, bv = (x >= a && x < ab) || x >= abc;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            src[i] = u;
            dst[i] = v;
        }
    }
    // seeded vertex permutation (Fisher-Yates on a separate stream)
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    Rng pr(chunk_seed(seed ^ 0x5045524d55544531ull, 0));
    for (uint64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[pr.below(i + 1)]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)m; i++) {
        src[i] = perm[src[i]];
        dst[i] = perm[dst[i]];
    }
    auto* g = new g2h_graph;
    g2h::csr_from_edges(n, src, dst, *g, omp_get_max_threads());
    *out = g;
    return 0;
}

int g2h_gen_social(uint64_t n, double gamma, double w_min, double w_max, uint64_t backbone_edges,
                   uint64_t ncomm, uint32_t smin, uint32_t smax, double comm_alpha,
                   double pmin, double pmax, double ecap, uint64_t seed, g2h_graph** out) {
    if (n < 2 || n > 0xffffffffull || smin < 2 || smax < smin || smax > n) {
        g2h::set_error("bad social-graph parameters");
        return -1;
    }
    auto w = chung_lu_weights(n, gamma, w_min, w_max, 2.0 * backbone_edges);
    const double sum = std::accumulate(w.begin(), w.end(), 0.0);
    const uint64_t mb = (uint64_t)std::llround(sum / 2.0);
    Alias al(w);
    // community sizes and densities (one stream, serial: ncomm draws)
    std::vector<double> sw(smax - smin + 1);
    for (uint32_t s = smin; s <= smax; s++) sw[s - smin] = std::pow((double)s, -comm_alpha);
    Alias sal(sw);
    Rng cr(chunk_seed(seed ^ 0x434f4d4d554e4954ull, 0));
    std::vector<uint32_t> csize(ncomm);
    std::vector<double> cp(ncomm);
    std::vector<uint64_t> coff(ncomm + 1, 0);
    // expected intra edges per community bound the output arrays
    for (uint64_t i = 0; i < ncomm; i++) {
        const uint32_t s = smin + sal.draw(cr);
        double p = pmin;
        for (int t = 0; t < 64; t++) {
            const double cand = pmin + (pmax - pmin) * cr.uniform();
            if (expected_max_cliques(s, cand) <= ecap) { p = cand; break; }
        }
        csize[i] = s;
        cp[i] = p;
        coff[i + 1] = coff[i] + (uint64_t)s * (s - 1) / 2;
    }
    const uint64_t pairs_max = coff[ncomm];
    std::vector<uint32_t> a(mb + pairs_max), b(mb + pairs_max);
    draw_pairs(al, mb, seed, a, b, 0);
    g2h::progress("social: %llu backbone pairs drawn", (unsigned long long)mb);
    std::vector<uint8_t> keep(pairs_max, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)ncomm; i++) {
        Rng r(chunk_seed(seed ^ 0x4d454d4245525321ull, i));
        const uint32_t s = csize[i];
        std::vector<uint32_t> mem;
        mem.reserve(s);
        while (mem.size() < s) {  // distinct uniform members
            const uint32_t v = (uint32_t)r.below(n);
            if (std::find(mem.begin(), mem.end(), v) == mem.end()) mem.push_back(v);
        }
        uint64_t k = coff[i];
        for (uint32_t x = 0; x < s; x++)
            for (uint32_t y = x + 1; y < s; y++, k++) {
                if (r.uniform() < cp[i]) {
                    a[mb + k] = mem[x];
                    b[mb + k] = mem[y];
                    keep[k] = 1;
                }
            }
    }
    // compact the kept community pairs behind the backbone draws
    uint64_t tot = mb;
    for (uint64_t k = 0; k < pairs_max; k++)
        if (keep[k]) { a[tot] = a[mb + k]; b[tot] = b[mb + k]; tot++; }
    a.resize(tot);
    b.resize(tot);
    g2h::progress("social: %llu community pairs kept", (unsigned long long)(tot - mb));
    auto* g = new g2h_graph;
    g2h::csr_from_edges(n, a, b, *g, omp_get_max_threads());
    g2h::progress("social: csr built, %llu entries", (unsigned long long)g->ne);
    *out = g;
    return 0;
}

}  // extern "C"
