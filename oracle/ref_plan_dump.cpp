// oracle/ref_plan_dump.cpp -- test program around the reference's own query-plan code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
// the reference's headers where they lie (common/graph.h, system/plan.h) and
// the vendored bliss sources, into oracle/_ref/plan_dump.  It prints the plan
// tables of a preset pattern or of an edge list as one JSON object, which
// tests/golden/make_plan_fixtures.py stores as golden fixtures for the
// product's own plan builder.  Nothing here ships or runs on the GPU.
//
// usage: plan_dump -q <preset id>
//        plan_dump -e <n> <u0> <v0> <u1> <v1> ...     (undirected edge list)
#include "common/graph.h"
#include "system/plan.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

static void arr(std::ostream& o, const char* k, const uintV* a, size_t n, bool comma = true) {
    o << "\"" << k << "\": [";
    for (size_t i = 0; i < n; i++) o << (i ? "," : "") << a[i];
    o << "]" << (comma ? ", " : "");
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: plan_dump -q <id> | -e <n> u v ..." << std::endl;
        return 2;
    }
    std::string mode = argv[1];
    Graph qg;
    // silence the reference's progress prints on stdout; JSON goes to fd 3 via a stream copy
    std::stringstream out;
    std::streambuf* saved = std::cout.rdbuf(std::cerr.rdbuf());
    if (mode == "-q") {
        qg = Graph("", (PresetPatternType)std::atoi(argv[2]), GraphType::QUERY);
    } else {
        const size_t n = std::strtoul(argv[2], nullptr, 10);
        std::vector<std::vector<uintV>> adj(n);
        for (int i = 3; i + 1 < argc; i += 2) {
            uintV a = std::atoi(argv[i]), b = std::atoi(argv[i + 1]);
            adj[a].push_back(b);
            adj[b].push_back(a);
        }
        size_t m = 0;
        for (auto& r : adj) { std::sort(r.begin(), r.end()); m += r.size(); }
        auto* rp = new uintE[n + 1];
        auto* cl = new uintV[m];
        rp[0] = 0;
        for (size_t v = 0; v < n; v++) {
            rp[v + 1] = rp[v] + adj[v].size();
            std::copy(adj[v].begin(), adj[v].end(), cl + rp[v]);
        }
        qg.SetVertexCount(n);
        qg.SetEdgeCount(m);
        qg.SetRowPtrs(rp);
        qg.SetCols(cl);
        qg.ReMapVertexId();
    }
    auto conds = qg.GetConditions(qg.GetBlissGraph());
    qg.SetConditions(conds);

    Plan plan;
    plan.graph = std::move(qg);
    plan.FindRoot();
    plan.GenerateSearchSequence();
    plan.GenerateBackwardNeighbor();
    plan.GeneratePreAfterBackwardNeighbor();
    plan.GenerateUsefulOrder();
    plan.GenerateStoreStrategy();
    std::cout.rdbuf(saved);

    const size_t n = plan.vertex_count_;
    const Graph& g = plan.graph;
    std::ostream& o = std::cout;
    o << "{";
    o << "\"n\": " << n << ", \"hop\": " << plan.GetHop() << ", \"root_degree\": " << plan.root_degree_ << ", ";
    // remapped query graph
    o << "\"rows\": [";
    for (size_t v = 0; v < n; v++) {
        o << (v ? "," : "") << "[";
        for (auto e = g.GetRowPtrs()[v]; e < g.GetRowPtrs()[v + 1]; e++)
            o << (e > g.GetRowPtrs()[v] ? "," : "") << g.GetCols()[e];
        o << "]";
    }
    o << "], ";
    o << "\"conditions\": [";
    for (size_t i = 0; i < conds.size(); i++) o << (i ? "," : "") << "[" << conds[i].first << "," << conds[i].second << "]";
    o << "], ";
    arr(o, "seq", plan.seq_.data(), n);
    arr(o, "rev", plan.reverse_seq_.data(), n);
    arr(o, "bn_count", plan.backNeighborCountHost, n);
    o << "\"bn\": [";
    for (size_t i = 0; i < n; i++) {
        o << (i ? "," : "") << "[";
        for (size_t j = 0; j < plan.backNeighborCountHost[i]; j++) o << (j ? "," : "") << plan.backNeighborsHost[i * n + j];
        o << "]";
    }
    o << "], ";
    arr(o, "cond_count", plan.condNumHost, n);
    auto condtab = [&](const char* key, const uintV* cnt, const uintV* tab) {
        o << "\"" << key << "\": [";
        for (size_t i = 0; i < n; i++) {
            o << (i ? "," : "") << "[";
            for (size_t k = 0; k < cnt[i]; k++)
                o << (k ? "," : "") << "[" << tab[2 * i * n + 2 * k] << "," << tab[2 * i * n + 2 * k + 1] << "]";
            o << "]";
        }
        o << "], ";
    };
    condtab("cond", plan.condNumHost, plan.condOrderHost);
    o << "\"shared\": [";
    for (size_t i = 0; i < n; i++) o << (i ? "," : "") << (plan.share_intersection[i] ? 1 : 0);
    o << "], ";
    arr(o, "pre_bn_count", plan.preBackNeighborCountHost, n);
    o << "\"pre_bn\": [";
    for (size_t i = 0; i < n; i++) {
        o << (i ? "," : "") << "[";
        for (size_t j = 0; j < plan.preBackNeighborCountHost[i]; j++) o << (j ? "," : "") << plan.preBackNeighborsHost[i * n + j];
        o << "]";
    }
    o << "], ";
    arr(o, "after_bn_count", plan.afterBackNeighborCountHost, n);
    o << "\"after_bn\": [";
    for (size_t i = 0; i < n; i++) {
        o << (i ? "," : "") << "[";
        for (size_t j = 0; j < plan.afterBackNeighborCountHost[i]; j++) o << (j ? "," : "") << plan.afterBackNeighborsHost[i * n + j];
        o << "]";
    }
    o << "], ";
    arr(o, "pre_cond_count", plan.preCondNumHost, n);
    condtab("pre_cond", plan.preCondNumHost, plan.preCondOrderHost);
    arr(o, "after_cond_count", plan.afterCondNumHost, n);
    condtab("after_cond", plan.afterCondNumHost, plan.afterCondOrderHost);
    o << "\"strategy\": [";
    for (size_t i = 0; i <= n; i++) o << (i ? "," : "") << (int)plan.strategy[i];
    o << "], ";
    arr(o, "moving_lvl", plan.moving_lvl, n, false);
    o << "}" << std::endl;
    return 0;
}
