// host/sched.cpp -- host side of the learned root scheduler (no counterpart
// in the reference, whose roots run in source order: app_BK/BK.h:323-355,
// system/pipeline_executor.h:151-207).
//
// The device computes five integer features per root (g2_features) and runs
// the MLP cost model (g2_mlp_*); this file turns them into the model's eight
// inputs, standardises them, turns predictions into a processing order and
// per-root split counts, computes core numbers, and reads / writes trained
// models.  Neither the order nor the split changes a result: they only
// decide which wave walks which part of which root's subtree, when.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.hpp"

extern "C" {

// Batagelj-Zaversnik bucket peel: repeatedly remove a vertex of smallest
// remaining degree; its core number is that degree at removal (kept
// monotone).  O(V + E).
int g2h_core_numbers(const uint64_t* rp, uint64_t nv, const uint32_t* cols, uint32_t* core) {
    if (!rp || !core || (rp[nv] && !cols)) { g2h::set_error("null argument"); return -1; }
    if (nv >= 0xffffffffull) { g2h::set_error("vertex ids must fit in u32"); return -1; }
    uint64_t md = 0;
    std::vector<uint32_t> deg(nv);
    for (uint64_t v = 0; v < nv; v++) {
        if (rp[v + 1] < rp[v]) { g2h::set_error("row_ptrs must be non-decreasing"); return -1; }
        deg[v] = (uint32_t)std::min<uint64_t>(rp[v + 1] - rp[v], 0xffffffffull);
        md = std::max<uint64_t>(md, deg[v]);
    }
    for (uint64_t j = rp[0]; j < rp[nv]; j++)
        if (cols[j] >= nv) { g2h::set_error("column id out of range"); return -1; }
    // bin[d] = first position of degree d in vert (vertices sorted by degree)
    std::vector<uint64_t> bin(md + 2, 0);
    for (uint64_t v = 0; v < nv; v++) bin[deg[v] + 1]++;
    for (uint64_t d = 1; d <= md + 1; d++) bin[d] += bin[d - 1];
    std::vector<uint32_t> vert(nv), pos(nv);
    {
        std::vector<uint64_t> next(bin.begin(), bin.end() - 1);
        for (uint64_t v = 0; v < nv; v++) {
            pos[v] = (uint32_t)next[deg[v]]++;
            vert[pos[v]] = (uint32_t)v;
        }
    }
    for (uint64_t i = 0; i < nv; i++) {
        const uint32_t v = vert[i];
        core[v] = deg[v];
        for (uint64_t j = rp[v]; j < rp[v + 1]; j++) {
            const uint32_t u = cols[j];
            if (deg[u] > deg[v]) {  // move u to the front of its bin, one degree down
                const uint32_t du = deg[u];
                const uint32_t pu = pos[u], pw = (uint32_t)bin[du];
                const uint32_t w = vert[pw];
                if (u != w) {
                    vert[pu] = w;
                    pos[w] = pu;
                    vert[pw] = u;
                    pos[u] = pw;
                }
                bin[du]++;
                deg[u]--;
            }
        }
    }
    return 0;
}

int g2h_root_features(const int64_t* feat, const uint32_t* sources, uint64_t ns, const uint32_t* core,
                      uint64_t nv, float* out) {
    if ((ns && (!feat || !sources || !core || !out))) { g2h::set_error("null argument"); return -1; }
    for (uint64_t i = 0; i < ns; i++) {
        if (sources[i] >= nv) { g2h::set_error("source id out of range"); return -1; }
        const int64_t* f = feat + i * G2H_NFEAT_INT;
        const double deg = (double)f[0], two = (double)f[1], tri = (double)f[2], shared = (double)f[3],
                     p0 = (double)f[4];
        const double pairs = deg * (deg - 1);
        float* x = out + i * G2H_NFEAT;
        x[0] = (float)std::log1p(deg);
        x[1] = (float)std::log1p((double)core[sources[i]]);
        x[2] = (float)std::log1p(two);
        x[3] = (float)std::log1p(tri);
        x[4] = (float)(pairs > 0 ? 2.0 * tri / pairs : 0.0);
        x[5] = (float)((double)sources[i] / (double)std::max<uint64_t>(nv, 1));
        x[6] = (float)(shared / std::max(deg, 1.0));
        x[7] = (float)std::log1p(p0);
    }
    return 0;
}

int g2h_norm_fit(const float* x, uint64_t n, g2h_norm* out) {
    if (!out || (n && !x)) { g2h::set_error("null argument"); return -1; }
    for (int k = 0; k < G2H_NFEAT; k++) {
        double s = 0, s2 = 0;
        for (uint64_t i = 0; i < n; i++) s += x[i * G2H_NFEAT + k];
        const double mu = n ? s / (double)n : 0.0;
        for (uint64_t i = 0; i < n; i++) {
            const double d = x[i * G2H_NFEAT + k] - mu;
            s2 += d * d;
        }
        const double sd = n ? std::sqrt(s2 / (double)n) : 0.0;
        out->mu[k] = (float)mu;
        out->sd[k] = (float)std::max(sd, 1e-6);
    }
    return 0;
}

void g2h_norm_apply(const g2h_norm* nrm, float* x, uint64_t n) {
    for (uint64_t i = 0; i < n; i++)
        for (int k = 0; k < G2H_NFEAT; k++)
            x[i * G2H_NFEAT + k] = (float)(((double)x[i * G2H_NFEAT + k] - nrm->mu[k]) / nrm->sd[k]);
}

// Longest predicted cost first; a root predicted above the per-wave share
// of the total is dealt into ceil(cost / share) parts, at most its number
// of later neighbours (the first-level children it can have) and max_split.
int g2h_learned_schedule(const float* pred, const int64_t* later, uint64_t ns, uint64_t waves,
                         uint32_t max_split, uint32_t* order, uint32_t* split) {
    if (ns && (!pred || !later || !order || !split)) { g2h::set_error("null argument"); return -1; }
    if (ns >= 0xffffffffull) { g2h::set_error("at most 2^32-1 roots"); return -1; }
    if (!max_split) max_split = 4096;
    std::vector<double> cost(ns);
    std::vector<uint64_t> a(ns), b(ns);  // ~bits(p) << 32 | index
    double sum = 0;
    for (uint64_t i = 0; i < ns; i++) {
        const float p = std::isfinite(pred[i]) ? std::min(std::max(pred[i], 0.0f), 60.0f) : 0.0f;
        cost[i] = std::expm1((double)p);
        sum += cost[i];
        uint32_t bits;  // a non-negative float's bits order like its value
        std::memcpy(&bits, &p, 4);
        a[i] = (uint64_t)~bits << 32 | i;
    }
    // longest predicted first (expm1 is increasing, so ~bits(p) ascending is
    // cost descending), ties in source order: a stable LSD radix sort on the
    // high 32 bits, 8 per pass, of keys that start in index order
    for (int shift = 32; shift < 64; shift += 8) {
        uint64_t cnt[257] = {0};
        for (uint64_t i = 0; i < ns; i++) cnt[((a[i] >> shift) & 255) + 1]++;
        for (int k = 0; k < 256; k++) cnt[k + 1] += cnt[k];
        for (uint64_t i = 0; i < ns; i++) b[cnt[(a[i] >> shift) & 255]++] = a[i];
        a.swap(b);
    }
    for (uint64_t i = 0; i < ns; i++) order[i] = (uint32_t)a[i];
    const double share = std::max(sum / (double)std::max<uint64_t>(waves, 1), 1.0);
    for (uint64_t i = 0; i < ns; i++) {
        double k = std::ceil(cost[i] / share);
        k = std::min(k, (double)std::max<int64_t>(later[i], 1));
        k = std::min(std::max(k, 1.0), (double)max_split);
        split[i] = (uint32_t)k;
    }
    return 0;
}

// Model file: "G2MLP1\n", then the eight means, the eight scales and the
// parameters, each one float per line in %.9g (round-trips exactly).
int g2h_model_save(const char* path, const g2h_norm* nrm, const float* params, uint64_t nparams) {
    if (!path || !nrm || (nparams && !params)) { g2h::set_error("null argument"); return -1; }
    FILE* f = std::fopen(path, "w");
    if (!f) { g2h::set_error(std::string("cannot write ") + path); return -2; }
    std::fprintf(f, "G2MLP1\n%d %llu\n", G2H_NFEAT, (unsigned long long)nparams);
    for (int k = 0; k < G2H_NFEAT; k++) std::fprintf(f, "%.9g\n", nrm->mu[k]);
    for (int k = 0; k < G2H_NFEAT; k++) std::fprintf(f, "%.9g\n", nrm->sd[k]);
    for (uint64_t i = 0; i < nparams; i++) std::fprintf(f, "%.9g\n", params[i]);
    const bool ok = std::fflush(f) == 0 && !std::ferror(f);
    std::fclose(f);
    if (!ok) { g2h::set_error(std::string("write error on ") + path); return -2; }
    return 0;
}

int g2h_model_load(const char* path, g2h_norm* nrm, float* params, uint64_t nparams) {
    if (!path || !nrm || (nparams && !params)) { g2h::set_error("null argument"); return -1; }
    FILE* f = std::fopen(path, "r");
    if (!f) { g2h::set_error(std::string("cannot read ") + path); return -2; }
    char magic[16] = {0};
    int nf = 0;
    unsigned long long np = 0;
    bool ok = std::fscanf(f, "%15s %d %llu", magic, &nf, &np) == 3 && std::strcmp(magic, "G2MLP1") == 0 &&
              nf == G2H_NFEAT && np == nparams;
    for (int k = 0; ok && k < G2H_NFEAT; k++) ok = std::fscanf(f, "%f", &nrm->mu[k]) == 1;
    for (int k = 0; ok && k < G2H_NFEAT; k++) ok = std::fscanf(f, "%f", &nrm->sd[k]) == 1 && nrm->sd[k] > 0;
    for (uint64_t i = 0; ok && i < nparams; i++) ok = std::fscanf(f, "%f", &params[i]) == 1 && std::isfinite(params[i]);
    std::fclose(f);
    if (!ok) { g2h::set_error(std::string("not a model file of this network: ") + path); return -3; }
    return 0;
}

}  // extern "C"
