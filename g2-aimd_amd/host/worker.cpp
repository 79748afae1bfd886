// host/worker.cpp -- Worker::run restated over libg2host.so / libg2hip.so.
#include "worker.hpp"

#include <algorithm>
#include <fstream>
#include <numeric>

namespace g2w {

// ---- bins (view/view_bin_manager.h) -----------------------------------------
std::vector<ViewBin> load_bins(const std::string& file, uint64_t nv, size_t* max_sources) {
    Timer t;
    t.start();
    std::vector<uint32_t> map(std::max<uint64_t>(nv, 1));
    uint32_t nb = 0;
    if (g2h_read_vbmap(file.c_str(), &nb, map.data(), nv) != 0)
        die(std::string("cannot load view bin partition file ") + file + ": " + g2h_last_error());
    t.stop();
    say(t.line("loaded view bin partition file, " + std::to_string(nb) + " view bins"));
    std::vector<ViewBin> pool(nb);
    for (uint32_t b = 0; b < nb; b++) pool[b].id = b;
    for (uint64_t v = 0; v < nv; v++) pool[map[v]].sources.push_back((uint32_t)v);
    *max_sources = 0;
    for (uint32_t b = 0; b < nb; b++) {
        say("Source " + std::to_string(b) + ": " + std::to_string(pool[b].sources.size()));
        *max_sources = std::max(*max_sources, pool[b].sources.size());
    }
    say("max partitioned sources num: " + std::to_string(*max_sources));
    return pool;
}

void reorder_bins(std::vector<ViewBin>& pool, const uint64_t* rp) {
    for (auto& b : pool) say(std::to_string(b.id) + " " + std::to_string(b.factor(rp)));
    std::stable_sort(pool.begin(), pool.end(),
                     [rp](const ViewBin& a, const ViewBin& b) { return a.factor(rp) > b.factor(rp); });
    for (auto& b : pool) say(std::to_string(b.id) + " " + std::to_string(b.factor(rp)));
}

// ViewBinManager::Split: split the bin furthest above the mean factor into two
// halves (alternating by degree) until there are (K/d + times)*d bins
void split_bins(std::vector<ViewBin>& pool, const uint64_t* rp, int device_num, int times) {
    int k = pool.size();
    const int target = (k / device_num + times) * device_num;
    int splits = target - k;
    while (splits-- > 0) {
        uint64_t sum = 0;
        for (auto& b : pool) sum += b.factor(rp);
        const uint64_t avg = sum / pool.size();
        uint64_t maxd = 0;
        size_t pos = 0;
        for (size_t i = 0; i < pool.size(); i++) {
            const uint64_t f = pool[i].factor(rp);
            if (f < avg) continue;
            if (f - avg > maxd) { maxd = f - avg; pos = i; }
        }
        std::vector<std::pair<uint32_t, uint64_t>> sf;
        for (uint32_t v : pool[pos].sources) sf.push_back({v, rp[v + 1] - rp[v]});
        std::stable_sort(sf.begin(), sf.end(), [](const auto& a, const auto& b) { return a.second > b.second; });
        ViewBin a, b;
        for (size_t i = 0; i < sf.size(); i++) (i % 2 ? a : b).sources.push_back(sf[i].first);
        const int removed = pool[pos].id;
        pool.erase(pool.begin() + pos);
        for (auto& x : pool)
            if (x.id > removed) x.id--;
        a.id = pool.size();
        b.id = pool.size() + 1;
        pool.push_back(std::move(a));
        pool.push_back(std::move(b));
    }
    std::stable_sort(pool.begin(), pool.end(),
                     [rp](const ViewBin& a, const ViewBin& b) { return a.factor(rp) > b.factor(rp); });
    for (auto& b : pool) say(std::to_string(b.id) + " " + std::to_string(b.factor(rp)));
}

// ---- executor ---------------------------------------------------------------
Executor::Executor(int consumer, int device, uint64_t budget) : consumer_(consumer) {
    tag_ = "device " + std::to_string(consumer);
    Timer t;
    t.start();
    g2_tuning tune{};
    const int rc = g2_create(&ctx_, device, budget, &tune);
    if (rc != 0) die("hipError: cannot create a device context on GPU " + std::to_string(device));
    t.stop();
    say(t.line(tag_ + " executor construction"));
}

Executor::~Executor() { g2_destroy(ctx_); }

void Executor::check(int rc, const char* what) {
    if (rc != 0) die(std::string(what) + ": " + g2_last_error(ctx_));
}

void Executor::transfer(const Holder& h) {
    say(tag_ + " transfer view bin " + std::to_string(h.view_id));
    Timer t;
    t.start();
    check(g2_upload_view(ctx_, h.rp, h.nv, h.cols, h.ne, h.sources.data(), h.sources.size()), "transfer");
    ns_ = h.sources.size();
    t.stop();
    say(t.line(tag_ + " memory copy"));
}

void Executor::set_plan(const g2_plan& p) { check(g2_gm_set_plan(ctx_, &p), "plan"); }

// PipelineExecutor::Run with the BK app: roots in AIMD-sized chunks; a chunk
// whose deferred large roots do not fit the memory budget is an overflow and
// is retried with a smaller chunk (system/pipeline_executor.h:151-207)
RunResult Executor::run_bk(uint32_t hist_len) {
    RunResult r;
    r.hist.assign(hist_len, 0);
    std::vector<uint64_t> h(hist_len);
    g2h_aimd aimd;
    g2h_aimd_init(&aimd);
    uint64_t total = 0;
    for (uint64_t i = 0; i < ns_;) {
        g2_sched s{nullptr, ns_, i, std::min<uint64_t>(ns_, i + aimd.chunk), nullptr};
        uint64_t cl = 0;
        const int rc = g2_mce_run(ctx_, &s, 0, &cl, h.data(), hist_len, nullptr, nullptr);
        bool overflow = false;
        if (rc == G2_EBUDGET) {
            if (aimd.chunk <= 1) die("Host overflow occured: " + std::string(g2_last_error(ctx_)));
            say("Iteration failed!!");
            overflow = true;
        } else {
            check(rc, "bk run");
            i = s.end;
            total += cl;
            for (uint32_t k = 0; k < hist_len; k++) r.hist[k] += h[k];
        }
        const char* tag = g2h_aimd_adjust(&aimd, overflow);
        say(std::string(tag) + std::to_string(aimd.chunk) + " cliques " + std::to_string(total));
    }
    say("Total No. of Cliques: " + std::to_string(total));
    r.count = total;
    return r;
}

RunResult Executor::run_gm() {
    RunResult r;
    check(g2_gm_run(ctx_, 0, &r.count, nullptr, nullptr), "gm run");
    return r;
}

// ---- plan printout (system/worker.h:62-154) ---------------------------------
static const char* op_name(uint32_t op) {
    return op == 0 ? "LESS_THAN" : op == 1 ? "LAGER_THAN" : op == 2 ? "NON_EQUAL" : "";
}

static void print_plan(const g2_plan& p, uint32_t n, const uint32_t* qadj) {
    uint32_t npairs = 0;
    uint32_t pairs[2 * G2_MAXQ * G2_MAXQ];
    g2h_query_conditions(n, qadj, &npairs, pairs);
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> order(n);
    for (uint32_t k = 0; k < npairs; k++) {
        order[pairs[2 * k]].push_back({0, pairs[2 * k + 1]});
        order[pairs[2 * k + 1]].push_back({1, pairs[2 * k]});
    }
    say("conditions: ");
    for (uint32_t i = 0; i < n; i++) {
        std::string s = std::to_string(i) + ": ";
        for (auto& c : order[i]) s += std::string(op_name(c.first)) + "(" + std::to_string(c.second) + "), ";
        say(s);
    }
    say("Hop = " + std::to_string(p.hop));
    say("Root degree = " + std::to_string(p.root_degree));
    std::string s = "Match Order: ";
    for (uint32_t i = 0; i < n; i++) s += std::to_string(p.seq[i]) + " ";
    say(s);
    s = "Moving Level: ";
    for (uint32_t i = 0; i < n; i++) s += std::to_string(p.moving_lvl[i]) + " ";
    say(s);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t pos = p.rev[i];
        s = std::to_string(i) + ": ";
        for (uint32_t j = 0; j < p.bn_count[pos]; j++) s += std::to_string(p.bn[pos][j]) + " ";
        say(s);
    }
    for (uint32_t i = 0; i < n; i++) {
        if (!p.shared[i]) continue;
        s = "shared loc = " + std::to_string(i) + "       pre BN: ";
        for (uint32_t j = 0; j < p.pre_bn_count[i]; j++) s += std::to_string(p.pre_bn[i][j]) + " ";
        s += "after BN: ";
        for (uint32_t j = 0; j < p.after_bn_count[i]; j++) s += std::to_string(p.after_bn[i][j]) + " ";
        say(s);
    }
}

// ---- Worker::run ------------------------------------------------------------
int run_worker(int argc, char** argv, App app) {
    CommandLine cmd{argc, argv};
    const std::string data = cmd.get("-dg", "./data/com-friendster.ungraph.txt.bin");
    const int query = cmd.get_int("-q", 1000);
    const double mem = cmd.get_double("-m", 10);
    const int threads = cmd.get_int("-t", 1);
    const int qs = cmd.get_int("-qs", 1);
    int hop = cmd.get_int("-h", -1);
    const int pn = cmd.get_int("-pn", 1);
    const int cn = cmd.get_int("-cn", 1);
    const int dr = cmd.get_int("-dr", 0);
    const int ds = cmd.get_int("-ds", 0);
    const int dst = cmd.get_int("-dst", 1);
    const int ss = cmd.get_int("-ss", 0);
    const double budget_gb = cmd.get_double("-budget", 0);  // per consumer, 0 = share the device evenly
    const std::string json = cmd.get("-json", "");
    (void)threads;

    const int gpus = g2_device_count();
    say("GPU count: " + std::to_string(gpus));
    if (gpus < 1) die("no GPU visible");
    if (pn < 1 || cn < 1) die("producer and consumer counts must be >= 1");
    if (qs < cn) die("queue size must be >= consumer count");
    {
        std::ostringstream o;
        o << "m: " << mem << " t: " << threads << " qs: " << qs << " pn: " << pn << " cn: " << cn;
        say(o.str());
    }

    Timer rt;
    rt.start();
    say("start reading CSR bin file....");
    g2h_graph* g = nullptr;
    if (g2h_graph_read_bin(data.c_str(), &g) != 0) die(std::string("cannot read ") + data + ": " + g2h_last_error());
    const uint64_t nv = g2h_graph_nv(g), ne = g2h_graph_ne(g);
    const uint64_t* rp = g2h_graph_row_ptrs(g);
    say("vertex_count=" + std::to_string(nv) + ", edge_count=" + std::to_string(ne));
    say("max degree=" + std::to_string(g2h_graph_max_degree(g)));
    rt.stop();
    say(rt.line("reading CSR bin file"));

    g2_plan plan{};
    uint64_t root_degree = 0;
    if (app == App::GM) {
        if (query == 1000) die("run needs -q <pattern>");
        if (hop != -1) die("-h is derived from the query plan for run");
        uint32_t n = 0, qadj[G2_MAXQ];
        if (g2h_query_preset(query, &n, qadj) != 0) die(std::string("query: ") + g2h_last_error());
        if (g2h_plan_build(n, qadj, &plan) != 0) die(std::string("plan: ") + g2h_last_error());
        print_plan(plan, n, qadj);
        hop = plan.hop;
        root_degree = plan.root_degree;
    }

    size_t max_sources = 0;
    auto pool = load_bins(data + "." + std::to_string(hop) + ".hop.vbmap", nv, &max_sources);
    if (ss)
        for (auto& b : pool) std::sort(b.sources.begin(), b.sources.end());
    if (dr) reorder_bins(pool, rp);
    if (ds) split_bins(pool, rp, cn, dst);
    for (size_t i = 0; i < pool.size(); i++) say(std::to_string(i) + " " + std::to_string(pool[i].id));

    // holders: a view never exceeds the whole graph
    const uint64_t cap = std::min<uint64_t>((uint64_t)(mem * 1e9) / 4, std::max<uint64_t>(ne, 1));
    Timer bt;
    bt.start();
    std::vector<std::unique_ptr<Holder>> holders;
    for (int i = 0; i < qs; i++) holders.emplace_back(new Holder(nv, cap));
    bt.stop();
    say(bt.line("view bin buffer"));

    BoundedQueue assigned(qs), released(qs);
    for (int i = 0; i < qs; i++) released.push(i);
    std::atomic<size_t> next_bin{0};
    std::atomic<int> finished{0};
    std::vector<std::thread> th;
    for (int p = 0; p < pn; p++) {
        th.emplace_back([&, p] {
            for (size_t b; (b = next_bin++) < pool.size();) {
                int h = -1;
                released.pop(h);
                say("producer " + std::to_string(p) + " gets " + std::to_string(h) + ", view_bin_id = " + std::to_string(b));
                Timer mt;
                mt.start();
                Holder& H = *holders[h];
                const auto& src = pool[b].sources;
                if (g2h_view_materialize(g, src.data(), src.size(), hop, root_degree, H.rp, H.cols, H.cap, &H.ne) != 0)
                    die(std::string("view bin materialize: ") + g2h_last_error());
                H.sources = src;
                H.view_id = pool[b].id;
                mt.stop();
                say(mt.line("view bin materialized " + std::to_string(pool[b].id)));
                say("producer " + std::to_string(p) + " produces " + std::to_string(b) + " in " + std::to_string(h));
                assigned.push(h);
            }
            finished++;
        });
    }

    // memory budget per consumer: the device's memory shared by its consumers
    std::vector<uint64_t> budget(cn, 0);
    for (int c = 0; c < cn; c++) {
        uint64_t tot = 0;
        g2_device_info(c % gpus, nullptr, &tot, nullptr, 0);
        const int sharing = (cn - c % gpus + gpus - 1) / gpus;
        budget[c] = budget_gb > 0 ? (uint64_t)(budget_gb * 1e9) : (uint64_t)(tot * 0.9) / std::max(1, sharing);
    }
    std::vector<RunResult> results(cn);
    for (int c = 0; c < cn; c++) {
        th.emplace_back([&, c] {
            Executor ex(c, c % gpus, budget[c]);
            if (app == App::GM) ex.set_plan(plan);
            RunResult acc;
            acc.hist.assign(128, 0);
            while (true) {
                int h = -1;
                if (finished == pn) {
                    if (!assigned.try_pop(h)) break;
                } else if (!assigned.try_pop(h)) {
                    std::this_thread::yield();
                    continue;
                }
                say("consumer " + std::to_string(c) + " gets " + std::to_string(h));
                ex.transfer(*holders[h]);
                say("consumer " + std::to_string(c) + " completes transfer " + std::to_string(h));
                released.push(h);
                Timer it;
                it.start();
                RunResult r = app == App::BK ? ex.run_bk(128) : ex.run_gm();
                it.stop();
                say(it.line("In-memory Execution Time"));
                say("consumer " + std::to_string(c) + " completes match " + std::to_string(h));
                acc.count += r.count;
                for (size_t k = 0; k < r.hist.size() && k < acc.hist.size(); k++) acc.hist[k] += r.hist[k];
                say("Currently Total counts: " + std::to_string(acc.count));
            }
            results[c] = acc;
            say("consumer " + std::to_string(c) + " stops");
        });
    }
    Timer tt;
    tt.start();
    for (auto& t : th) t.join();
    tt.stop();
    say(tt.line("total time"));
    uint64_t total = 0;
    std::vector<uint64_t> hist(128, 0);
    for (auto& r : results) {
        total += r.count;
        for (size_t k = 0; k < hist.size(); k++) hist[k] += r.hist[k];
    }
    say("Total counts: " + std::to_string(total));
    if (app == App::BK) {
        std::string s = "Clique size histogram:";
        for (size_t k = 0; k < hist.size(); k++)
            if (hist[k]) s += " " + std::to_string(k) + ":" + std::to_string(hist[k]);
        say(s);
    }
    if (!json.empty()) {
        std::ofstream f(json);
        f << "{\"app\": \"" << (app == App::BK ? "bk" : "run") << "\", \"graph\": \"" << data
          << "\", \"query\": " << query << ", \"total\": " << total << ", \"seconds\": " << tt.ms() / 1e3
          << ", \"hist\": {";
        bool first = true;
        for (size_t k = 0; k < hist.size(); k++)
            if (hist[k]) {
                f << (first ? "" : ", ") << "\"" << k << "\": " << hist[k];
                first = false;
            }
        f << "}}\n";
    }
    g2h_graph_free(g);
    return 0;
}

}  // namespace g2w
