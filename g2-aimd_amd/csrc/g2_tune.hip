// csrc/g2_tune.hip -- autotuner of the batched matching launch (host code).
//
// The batched launch has three shape parameters whose best values depend on
// the graph's degree distribution and the query mix: the level-1 window (how
// finely a hub root's row is cut into tasks), the root tasks a wave takes per
// fetch (fetch atomics vs tail granularity) and the resident waves per CU
// (latency hiding vs arena memory and L2 pressure).  The tuner times the real
// launch on a seeded sample of the queries and keeps the fastest shape,
// sweeping one parameter at a time from a fixed start point.  Counts do not
// depend on the shape, so every trial's counts are also checked against the
// first trial's.
#include <algorithm>
#include <numeric>
#include <random>

#include "g2_internal.hpp"

using namespace g2;

extern "C" int g2_gm_autotune(g2_ctx* c, const g2_plan* plans, uint32_t nq, uint32_t sample_queries,
                              uint64_t seed, g2_batch* best, double* best_ms) {
    if (!c || !best) return G2_EINVAL;
    if (nq == 0 || !plans) return set_err(c, G2_EINVAL, "no queries to tune on");
    const uint64_t budget = best->node_budget;
    std::vector<uint32_t> pick(nq);
    std::iota(pick.begin(), pick.end(), 0u);
    if (sample_queries && sample_queries < nq) {
        std::mt19937_64 rng(seed);
        for (uint32_t i = 0; i < sample_queries; i++) {
            std::uniform_int_distribution<uint32_t> d(i, nq - 1);
            std::swap(pick[i], pick[d(rng)]);
        }
        pick.resize(sample_queries);
        std::sort(pick.begin(), pick.end());
    }
    std::vector<g2_plan> sample;
    for (uint32_t q : pick) sample.push_back(plans[q]);

    g2_batch cur{};
    cur.window = 256;
    cur.roots_per_task = 1;
    cur.waves_per_cu = waves_per_cu(c, WK_GM);
    cur.node_budget = budget;
    std::vector<uint64_t> ref(sample.size()), got(sample.size());
    std::vector<uint32_t> over(sample.size());
    g2_stats st{};
    int rc = g2_gm_run_batch(c, sample.data(), (uint32_t)sample.size(), &cur, ref.data(), nullptr, over.data(), &st);
    if (rc) return rc;
    // queries over the budget are left out of the timing
    std::vector<g2_plan> kept;
    std::vector<uint64_t> kref;
    for (size_t i = 0; i < sample.size(); i++)
        if (!over[i]) {
            kept.push_back(sample[i]);
            kref.push_back(ref[i]);
        }
    if (kept.empty()) return set_err(c, G2_EINVAL, "every sampled query passes the node budget");
    got.resize(kept.size());
    auto trial = [&](const g2_batch& b, double* ms) -> int {
        g2_stats s{};
        int r = g2_gm_run_batch(c, kept.data(), (uint32_t)kept.size(), &b, got.data(), nullptr, nullptr, &s);
        if (r) return r;
        if (got != kref) return set_err(c, G2_EINVAL, "batched counts changed with the launch shape");
        *ms = s.kernel_ms;
        return G2_OK;
    };
    double cur_ms = 0;
    if ((rc = trial(cur, &cur_ms))) return rc;  // the first timed trial also warms caches
    if ((rc = trial(cur, &cur_ms))) return rc;
    const uint32_t windows[] = {0, 64, 256, 1024, 4096};
    const uint32_t rpts[] = {1, 2, 8, 32};
    const uint32_t wpcs[] = {8, 16, 24, 32};
    for (uint32_t w : windows) {
        if (w == cur.window) continue;
        g2_batch b = cur;
        b.window = w;
        double ms;
        if ((rc = trial(b, &ms))) return rc;
        if (ms < cur_ms) { cur = b; cur_ms = ms; }
    }
    for (uint32_t r : rpts) {
        if (r == cur.roots_per_task) continue;
        g2_batch b = cur;
        b.roots_per_task = r;
        double ms;
        if ((rc = trial(b, &ms))) return rc;
        if (ms < cur_ms) { cur = b; cur_ms = ms; }
    }
    for (uint32_t w : wpcs) {
        if (w == cur.waves_per_cu) continue;
        g2_batch b = cur;
        b.waves_per_cu = w;
        double ms;
        if ((rc = trial(b, &ms))) return rc;
        if (ms < cur_ms) { cur = b; cur_ms = ms; }
    }
    cur.order = nullptr;
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return G2_OK;
}
