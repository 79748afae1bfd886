// host/worker.cpp -- Worker::run restated over libg2host.so / libg2hip.so.
#include "worker.hpp"

#include <iomanip>
#include <algorithm>
#include <cmath>
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
void RunResult::add(const RunResult& r) {
    count += r.count;
    if (hist.size() < r.hist.size()) hist.resize(r.hist.size(), 0);
    for (size_t k = 0; k < r.hist.size(); k++) hist[k] += r.hist[k];
    kernel_ms += r.kernel_ms;
    sched_ms += r.sched_ms;
    idle_sum += r.idle_sum;
    launches += r.launches;
    retries += r.retries;
    nodes += r.nodes;
    donations += r.donations;
    split_roots += r.split_roots;
    local_views += r.local_views;
    spans.insert(spans.end(), r.spans.begin(), r.spans.end());
    host_spill_bytes = std::max(host_spill_bytes, r.host_spill_bytes);
    device_bytes = std::max(device_bytes, r.device_bytes);
}

Executor::Executor(int consumer, int device, uint64_t budget, const g2_tuning& tune)
    : consumer_(consumer), device_(device) {
    tag_ = "device " + std::to_string(consumer);
    Timer t;
    t.start();
    const int rc = g2_create(&ctx_, device, budget, &tune);
    if (rc != 0) die("hipError: cannot create a device context on GPU " + std::to_string(device));
    t.stop();
    say(t.line(tag_ + " executor construction"));
}

Executor::~Executor() {
    if (mlp_) g2_mlp_destroy(mlp_);
    g2_destroy(ctx_);
}

void Executor::check(int rc, const char* what) {
    if (rc != 0) die(std::string(what) + ": " + g2_last_error(ctx_));
}

void Executor::transfer(const Holder& h) {
    say(tag_ + " transfer view bin " + std::to_string(h.view_id));
    Timer t;
    t.start();
    check(g2_upload_view(ctx_, h.rp, h.nv, h.cols, h.ne, h.sources.data(), h.sources.size()), "transfer");
    ns_ = h.sources.size();
    sources_ = h.sources;
    t.stop();
    say(t.line(tag_ + " memory copy"));
}

void Executor::set_plan(const g2_plan& p) { check(g2_gm_set_plan(ctx_, &p), "plan"); }

// PipelineExecutor::Run with the BK app.  The view's roots run in one
// launch, in the form's order.  A launch the memory budget cannot hold is an
// overflow: the roots are then taken in AIMD-sized chunks of that order,
// retried smaller on every further overflow (system/pipeline_executor.h:151-207,
// system/subgraph_container.h:178-211).
RunResult Executor::run_bk(uint32_t hist_len, const BkForm& form) {
    RunResult r;
    r.hist.assign(hist_len, 0);
    std::vector<uint64_t> h(hist_len);
    std::vector<uint32_t> order, split;
    g2_sched sched{nullptr, ns_, 0, 0, nullptr};
    if (form.learned && ns_) {
        // cost model: device features -> 8 inputs -> standardise -> MLP on the
        // matrix cores -> longest predicted first, heavy roots split
        Timer st;
        st.start();
        if (!mlp_) {
            if (g2_mlp_create(&mlp_, device_, 1) != 0) die("cannot create the cost model on the device");
            if (g2_mlp_set_params(mlp_, form.model->params.data()) != 0) die("cannot load the cost model");
        }
        // features, inputs, standardisation, the cost model, the order and
        // the parts, all on the device; only order and parts come back
        const double t_sched = Timer::now_us();
        uint32_t cus = 0;
        g2_device_info(device_, &cus, nullptr, nullptr, 0);
        const uint64_t waves = (uint64_t)cus * (form.tune.waves_per_cu ? form.tune.waves_per_cu : 16);
        order.resize(ns_);
        split.resize(ns_);
        // with work donation on, roots stay whole (order only): donation
        // divides a heavy root where its work is, first-level parts each
        // rebuild the root's bitsets
        const uint32_t max_split = form.tune.mce_donate_every == 0xffffffffu ? 0u : 1u;
        check(g2_learned_schedule(ctx_, mlp_, form.model->norm.mu, form.model->norm.sd, form.core, waves, max_split,
                                  order.data(), split.data(), nullptr),
              "learned schedule");
        for (uint32_t k : split) r.split_roots += k > 1;
        sched.order = order.data();
        sched.split = split.data();
        st.stop();
        r.sched_ms = st.ms();
        std::ostringstream o;
        o << std::fixed << std::setprecision(1) << tag_ << ": schedule " << r.sched_ms << " ms = model set-up "
          << (t_sched - st.t0) / 1e3 << " + device schedule " << (st.t1 - t_sched) / 1e3 << " ms";
        say(o.str());
    }
    if (form.reverse && ns_) {
        // the sources backwards: in the degeneracy order the roots with the
        // largest neighbourhoods come last and leave a tail of long roots;
        // taken first, the short roots fill in behind them
        order.resize(ns_);
        for (uint64_t i = 0; i < ns_; i++) order[i] = (uint32_t)(ns_ - 1 - i);
        sched.order = order.data();
    }
    auto account = [&](const g2_stats& st, uint64_t cl) {
        r.count += cl;
        for (uint32_t k = 0; k < hist_len; k++) r.hist[k] += h[k];
        r.kernel_ms += st.kernel_ms;
        r.idle_sum += st.wave_idle_frac;
        r.launches++;
        r.nodes += st.nodes;
        r.donations += st.donations;
        r.host_spill_bytes = std::max(r.host_spill_bytes, st.host_spill_bytes);
        r.device_bytes = std::max(r.device_bytes, st.device_bytes);
    };
    g2_stats st{};
    uint64_t cl = 0;
    int rc = g2_mce_run(ctx_, form.learned || form.reverse ? &sched : nullptr, 0, &cl, h.data(), hist_len, nullptr, &st);
    if (rc == G2_OK) {
        account(st, cl);
    } else if (rc == G2_EBUDGET && ns_ > 1) {
        say("Iteration failed!! " + std::string(g2_last_error(ctx_)));
        r.retries++;
        g2h_aimd aimd;
        g2h_aimd_init(&aimd);
        const char* tag0 = g2h_aimd_adjust(&aimd, 1);
        say(std::string(tag0) + std::to_string(aimd.chunk));
        for (uint64_t i = 0; i < ns_;) {
            g2_sched s = sched;
            s.begin = i;
            s.end = std::min<uint64_t>(ns_, i + aimd.chunk);
            rc = g2_mce_run(ctx_, &s, 0, &cl, h.data(), hist_len, nullptr, &st);
            bool overflow = false;
            if (rc == G2_EBUDGET) {
                if (aimd.chunk <= 1) die("Host overflow occured: " + std::string(g2_last_error(ctx_)));
                say("Iteration failed!!");
                r.retries++;
                overflow = true;
            } else {
                check(rc, "bk run");
                i = s.end;
                account(st, cl);
            }
            const char* tag = g2h_aimd_adjust(&aimd, overflow);
            say(std::string(tag) + std::to_string(aimd.chunk) + " cliques " + std::to_string(r.count));
        }
    } else {
        check(rc, "bk run");
    }
    say("Total No. of Cliques: " + std::to_string(r.count));
    std::ostringstream o;
    o << tag_ << ": kernel " << r.kernel_ms << " ms, schedule " << r.sched_ms << " ms, wave idle "
      << (r.launches ? r.idle_sum / r.launches : 0.0) << ", donations " << r.donations << ", split roots "
      << r.split_roots << ", host spill " << r.host_spill_bytes << " B, device memory " << r.device_bytes << " B";
    say(o.str());
    return r;
}

RunResult Executor::run_gm() {
    RunResult r;
    g2_stats st{};
    check(g2_gm_run(ctx_, 0, &r.count, nullptr, &st), "gm run");
    r.kernel_ms = st.kernel_ms;
    r.launches = 1;
    r.nodes = st.nodes;
    r.device_bytes = st.device_bytes;
    r.local_views = st.local_bitmap;
    return r;
}

// ---- synthetic inputs (-gen) and cost-model training (-train) --------------
static void split_spec(const std::string& spec, std::string* kind, uint64_t* size, uint64_t* seed, uint64_t def_seed) {
    std::vector<std::string> part;
    std::stringstream ss(spec);
    for (std::string t; std::getline(ss, t, ':');) part.push_back(t);
    if (part.size() < 2 || part.size() > 3) die("graph spec must be kind:size[:seed], got " + spec);
    *kind = part[0];
    *size = std::strtoull(part[1].c_str(), nullptr, 0);
    *seed = part.size() == 3 ? std::strtoull(part[2].c_str(), nullptr, 0) : def_seed;
    if (!*size) die("graph spec with size 0: " + spec);
}

GenGraph generate(const std::string& spec) {
    std::string kind;
    uint64_t size = 0, seed = 0;
    GenGraph out;
    Timer t;
    t.start();
    int rc = -1;
    if (spec.rfind("social:", 0) == 0) {
        // social-network shape: Chung-Lu backbone (exponent 2.1, weights
        // 2..5000) plus planted communities of 10-200 vertices; 1.8e9
        // undirected edges at 65M vertices
        split_spec(spec, &kind, &size, &seed, 0x6732A1D0C0000003ull);
        const double f = (double)size / 65e6;
        rc = g2h_gen_social(size, 2.1, 2.0, 5000.0, (uint64_t)(1.54e9 * f), (uint64_t)(2e6 * f), 10, 200, 3.0, 0.3,
                            0.9, 1e4, seed, &out.g);
    } else if (spec.rfind("rmat:", 0) == 0) {
        split_spec(spec, &kind, &size, &seed, 0x6732A1D0C0000002ull);
        rc = g2h_gen_rmat((uint32_t)size, 16, 0.57, 0.19, 0.19, seed, &out.g);
    } else if (spec.rfind("chunglu:", 0) == 0) {
        split_spec(spec, &kind, &size, &seed, 0x6732A1D0C0000001ull);
        rc = g2h_gen_chung_lu(size, 2.3, 2.0, 2000.0, 10 * size, seed, &out.g);
    } else {
        die("unknown graph spec " + spec + " (social:<n>, rmat:<scale>, chunglu:<n>)");
    }
    if (rc != 0) die("generator: " + std::string(g2h_last_error()));
    if (g2h_graph_preprocess(out.g) != 0) die("preprocess: " + std::string(g2h_last_error()));
    const uint64_t nv = g2h_graph_nv(out.g);
    out.shells.assign(nv + 2, 0);
    uint64_t nsh = 0;
    uint32_t dg = 0;
    if (g2h_graph_kcore_recode(out.g, nullptr, out.shells.data(), &nsh, &dg) != 0)
        die("kcore recode: " + std::string(g2h_last_error()));
    out.shells.resize(nsh + 1);
    t.stop();
    say(t.line("generated " + spec + ": vertex_count=" + std::to_string(nv) + ", edge_count=" +
               std::to_string(g2h_graph_ne(out.g)) + ", degeneracy=" + std::to_string(dg)));
    return out;
}

Model train_model(const std::vector<std::string>& specs, int device, uint32_t epochs) {
    std::vector<float> X;
    std::vector<float> Y;
    for (const auto& spec : specs) {
        GenGraph gg = generate(spec);
        const uint64_t nv = g2h_graph_nv(gg.g), ne = g2h_graph_ne(gg.g);
        const uint64_t* rp = g2h_graph_row_ptrs(gg.g);
        std::vector<uint32_t> core(nv), src(nv);
        if (g2h_core_numbers(rp, nv, g2h_graph_cols(gg.g), core.data()) != 0)
            die("core numbers: " + std::string(g2h_last_error()));
        for (uint64_t v = 0; v < nv; v++) src[v] = (uint32_t)v;
        g2_ctx* c = nullptr;
        g2_tuning tune{};
        if (g2_create(&c, device, 0, &tune) != 0) die("cannot create a device context for training");
        auto ok = [&](int rc, const char* what) {
            if (rc != 0) die(std::string(what) + ": " + g2_last_error(c));
        };
        ok(g2_upload_view(c, rp, nv, g2h_graph_cols(gg.g), ne, src.data(), nv), "training upload");
        std::vector<int64_t> feat(nv * G2_NFEAT_INT);
        ok(g2_features(c, feat.data()), "training features");
        std::vector<uint64_t> nodes(nv), hist(128);
        uint64_t total = 0;
        ok(g2_mce_run(c, nullptr, 0, &total, hist.data(), 128, nodes.data(), nullptr), "training run");
        g2_destroy(c);
        const size_t base = X.size() / G2H_NFEAT;
        X.resize((base + nv) * G2H_NFEAT);
        Y.resize(base + nv);
        if (g2h_root_features(feat.data(), src.data(), nv, core.data(), nv, X.data() + base * G2H_NFEAT) != 0)
            die("training features: " + std::string(g2h_last_error()));
        for (uint64_t v = 0; v < nv; v++) Y[base + v] = (float)std::log1p((double)nodes[v]);
        say("training graph " + spec + ": " + std::to_string(nv) + " roots logged, " + std::to_string(total) +
            " maximal cliques");
        g2h_graph_free(gg.g);
    }
    Model m;
    const uint64_t n = Y.size();
    if (g2h_norm_fit(X.data(), n, &m.norm) != 0) die("standardiser: " + std::string(g2h_last_error()));
    g2h_norm_apply(&m.norm, X.data(), n);
    g2_mlp* mlp = nullptr;
    if (g2_mlp_create(&mlp, device, 1) != 0) die("cannot create the cost model");
    float loss = 0;
    if (g2_mlp_train(mlp, X.data(), Y.data(), n, epochs, 1024, 2e-3f, &loss) != 0) die("cost model training failed");
    m.params.resize(G2_MLP_NPARAM);
    g2_mlp_get_params(mlp, m.params.data());
    g2_mlp_destroy(mlp);
    std::ostringstream o;
    o << "cost model: " << n << " training roots, last-epoch mse " << loss;
    say(o.str());
    return m;
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
namespace {

struct PipelineResult {
    RunResult total;
    double seconds = 0;           // first materialisation to the last consumer's end
    std::vector<RunResult> per_consumer;
};

struct Pipeline {
    App app;
    const g2h_graph* g;
    std::vector<ViewBin>* pool;
    int hop;
    uint64_t root_degree;
    int pn, cn, qs, gpus;
    uint64_t cap;
    std::vector<uint64_t> budget;
    const g2_plan* plan;
    BkForm form;

    PipelineResult run() {
        PipelineResult out;
        std::vector<std::unique_ptr<Holder>> holders;
        {
            Timer bt;
            bt.start();
            const uint64_t nv = g2h_graph_nv(g);
            for (int i = 0; i < qs; i++) holders.emplace_back(new Holder(nv, cap));
            bt.stop();
            say(bt.line("view bin buffer"));
        }
        BoundedQueue assigned(qs), released(qs);
        for (int i = 0; i < qs; i++) released.push(i);
        std::atomic<size_t> next_bin{0};
        std::atomic<int> finished{0};
        std::vector<std::thread> th;
        for (int p = 0; p < pn; p++) {
            th.emplace_back([&, p] {
                for (size_t b; (b = next_bin++) < pool->size();) {
                    int h = -1;
                    released.pop(h);
                    say("producer " + std::to_string(p) + " gets " + std::to_string(h) + ", view_bin_id = " + std::to_string(b));
                    Timer mt;
                    mt.start();
                    Holder& H = *holders[h];
                    const auto& src = (*pool)[b].sources;
                    if (g2h_view_materialize(g, src.data(), src.size(), hop, root_degree, H.rp, H.cols, H.cap, &H.ne) != 0)
                        die(std::string("view bin materialize: ") + g2h_last_error());
                    H.sources = src;
                    H.view_id = (*pool)[b].id;
                    mt.stop();
                    say(mt.line("view bin materialized " + std::to_string((*pool)[b].id)));
                    say("producer " + std::to_string(p) + " produces " + std::to_string(b) + " in " + std::to_string(h));
                    assigned.push(h);
                }
                finished++;
            });
        }
        out.per_consumer.resize(cn);
        for (int c = 0; c < cn; c++) {
            th.emplace_back([&, c] {
                Executor ex(c, c % gpus, budget[c], form.tune);
                if (app == App::GM) ex.set_plan(*plan);
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
                    RunResult r = app == App::BK ? ex.run_bk(128, form) : ex.run_gm();
                    it.stop();
                    r.spans.push_back({it.t0, it.t1});
                    say(it.line("In-memory Execution Time"));
                    say("consumer " + std::to_string(c) + " completes match " + std::to_string(h));
                    acc.add(r);
                    say("Currently Total counts: " + std::to_string(acc.count));
                }
                out.per_consumer[c] = acc;
                say("consumer " + std::to_string(c) + " stops");
            });
        }
        Timer tt;
        tt.start();
        for (auto& t : th) t.join();
        tt.stop();
        say(tt.line("total time"));
        out.seconds = tt.ms() / 1e3;
        out.total.hist.assign(128, 0);
        for (auto& r : out.per_consumer) out.total.add(r);
        return out;
    }
};

void json_hist(std::ostream& f, const std::vector<uint64_t>& hist) {
    f << "{";
    bool first = true;
    for (size_t k = 0; k < hist.size(); k++)
        if (hist[k]) {
            f << (first ? "" : ", ") << "\"" << k << "\": " << hist[k];
            first = false;
        }
    f << "}";
}

// seconds during which at least one consumer was running a view (the union
// of the runs' wall-clock spans: device work plus the host schedule, without
// materialisation and copies)
double busy_seconds(std::vector<std::pair<double, double>> v) {
    std::sort(v.begin(), v.end());
    double tot = 0, a = 0, b = -1;
    for (auto& x : v) {
        if (x.first > b) {
            if (b > a) tot += b - a;
            a = x.first;
            b = x.second;
        } else {
            b = std::max(b, x.second);
        }
    }
    if (b > a) tot += b - a;
    return tot / 1e6;
}

void json_form(std::ostream& f, const PipelineResult& r) {
    const RunResult& t = r.total;
    f << "{\"total\": " << t.count << ", \"seconds\": " << r.seconds << ", \"run_seconds\": " << busy_seconds(t.spans)
      << ", \"kernel_ms\": " << t.kernel_ms
      << ", \"schedule_ms\": " << t.sched_ms << ", \"launches\": " << t.launches << ", \"retries\": " << t.retries
      << ", \"wave_idle_frac\": " << (t.launches ? t.idle_sum / t.launches : 0.0) << ", \"nodes\": " << t.nodes
      << ", \"donations\": " << t.donations << ", \"split_roots\": " << t.split_roots
      << ", \"host_spill_bytes\": " << t.host_spill_bytes << ", \"device_bytes\": " << t.device_bytes
      << ", \"consumer_kernel_ms\": [";
    for (size_t c = 0; c < r.per_consumer.size(); c++) f << (c ? ", " : "") << r.per_consumer[c].kernel_ms;
    f << "], \"hist\": ";
    json_hist(f, t.hist);
    f << "}";
}

// -qf: a pattern as text, one edge "a b" per line (ids from 0, at most 8
// vertices, connected; lines not starting with a digit are skipped)
void read_pattern(const std::string& path, uint32_t* n, uint32_t qadj[G2_MAXQ]) {
    std::ifstream in(path);
    if (!in) die("cannot read pattern file " + path);
    std::string line;
    uint32_t nv = 0;
    for (int k = 0; k < G2_MAXQ; k++) qadj[k] = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] < '0' || line[0] > '9') continue;
        std::istringstream ss(line);
        long a = -1, b = -1;
        if (!(ss >> a >> b) || a < 0 || b < 0 || a >= G2_MAXQ || b >= G2_MAXQ || a == b)
            die("pattern file " + path + ": bad edge line '" + line + "' (ids 0.." + std::to_string(G2_MAXQ - 1) + ")");
        qadj[a] |= 1u << b;
        qadj[b] |= 1u << a;
        nv = std::max<uint32_t>(nv, (uint32_t)std::max(a, b) + 1);
    }
    if (nv < 2) die("pattern file " + path + " has no edge");
    uint32_t seen = 1, frontier = 1;  // connected from vertex 0
    while (frontier) {
        uint32_t next = 0;
        for (uint32_t v = 0; v < nv; v++)
            if (frontier >> v & 1) next |= qadj[v];
        frontier = next & ~seen;
        seen |= next;
    }
    if (seen != (1u << nv) - 1) die("pattern file " + path + ": the pattern is not connected");
    *n = nv;
}

std::vector<std::string> split_list(const std::string& s) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    for (std::string t; std::getline(ss, t, ',');)
        if (!t.empty()) out.push_back(t);
    return out;
}

}  // namespace

int run_worker(int argc, char** argv, App app) {
    CommandLine cmd{argc, argv};
    const std::string data = cmd.get("-dg", "./data/com-friendster.ungraph.txt.bin");
    const std::string gen = cmd.get("-gen", "");
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
    const int nb = cmd.get_int("-nb", 4);
    const std::string bin_kind = cmd.get("-bins", "kcore");
    const double budget_gb = cmd.get_double("-budget", 0);  // per consumer, 0 = share the device evenly
    const std::string json = cmd.get("-json", "");
    const std::string query_file = cmd.get("-qf", "");
    const std::string sched = cmd.get("-sched", "plain");
    const std::string model_in = cmd.get("-model", "");
    const std::string train_out = cmd.get("-train", "");
    const std::string train_gen = cmd.get("-train-gen", "");
    const int epochs = cmd.get_int("-epochs", 20);
    const int rpw = cmd.get_int("-rpw", 2);
    const int wpc = cmd.get_int("-wpc", 24);
    const int donate = cmd.get_int("-donate", 0);
    (void)threads;

    g2h_set_progress(cmd.get_int("-progress", 0));  // progress lines of long host steps on stderr
    const int gpus = g2_device_count();
    say("GPU count: " + std::to_string(gpus));
    if (gpus < 1) die("no GPU visible");
    if (pn < 1 || cn < 1) die("producer and consumer counts must be >= 1");
    if (qs < cn) die("queue size must be >= consumer count");
    if (sched != "plain" && sched != "learned" && sched != "both" && sched != "reverse")
        die("-sched must be plain, learned, both or reverse");
    if (app == App::GM && sched != "plain") die("-sched applies to bk only");
    const bool want_learned = sched == "learned" || sched == "both";
    if (want_learned && model_in.empty() && train_out.empty()) die("-sched " + sched + " needs -model or -train");
    if (!train_out.empty() && train_gen.empty()) die("-train needs -train-gen <spec>[,<spec>...]");
    if (nb < 1) die("-nb must be >= 1");
    if (bin_kind != "kcore" && bin_kind != "kmeans") die("-bins must be kcore or kmeans");
    {
        std::ostringstream o;
        o << "m: " << mem << " t: " << threads << " qs: " << qs << " pn: " << pn << " cn: " << cn;
        say(o.str());
    }

    // the cost model first, so training graphs are freed before the input is built
    Model model;
    if (!train_out.empty()) {
        model = train_model(split_list(train_gen), 0, (uint32_t)epochs);
        if (g2h_model_save(train_out.c_str(), &model.norm, model.params.data(), model.params.size()) != 0)
            die("cannot write the cost model: " + std::string(g2h_last_error()));
        say("cost model written to " + train_out);
    } else if (!model_in.empty()) {
        model.params.resize(G2_MLP_NPARAM);
        if (g2h_model_load(model_in.c_str(), &model.norm, model.params.data(), model.params.size()) != 0)
            die("cannot load the cost model: " + std::string(g2h_last_error()));
    }

    Timer rt;
    rt.start();
    g2h_graph* g = nullptr;
    std::vector<uint64_t> shells;
    if (!gen.empty()) {
        GenGraph gg = generate(gen);
        g = gg.g;
        shells.swap(gg.shells);
    } else {
        say("start reading CSR bin file....");
        if (g2h_graph_read_bin(data.c_str(), &g) != 0) die(std::string("cannot read ") + data + ": " + g2h_last_error());
    }
    const uint64_t nv = g2h_graph_nv(g), ne = g2h_graph_ne(g);
    const uint64_t* rp = g2h_graph_row_ptrs(g);
    say("vertex_count=" + std::to_string(nv) + ", edge_count=" + std::to_string(ne));
    say("max degree=" + std::to_string(g2h_graph_max_degree(g)));
    rt.stop();
    say(rt.line(gen.empty() ? "reading CSR bin file" : "generating the graph"));

    g2_plan plan{};
    uint64_t root_degree = 0;
    if (app == App::GM) {
        if (query == 1000 && query_file.empty()) die("run needs -q <pattern> or -qf <pattern file>");
        if (hop != -1) die("-h is derived from the query plan for run");
        uint32_t n = 0, qadj[G2_MAXQ];
        if (!query_file.empty()) {
            uint32_t raw[G2_MAXQ] = {0};
            read_pattern(query_file, &n, raw);
            if (g2h_query_remap(n, raw, qadj) != 0) die(std::string("pattern: ") + g2h_last_error());
        } else if (g2h_query_preset(query, &n, qadj) != 0) {
            die(std::string("query: ") + g2h_last_error());
        }
        if (g2h_plan_build(n, qadj, &plan) != 0) die(std::string("plan: ") + g2h_last_error());
        print_plan(plan, n, qadj);
        hop = plan.hop;
        root_degree = plan.root_degree;
    }

    size_t max_sources = 0;
    std::vector<ViewBin> pool;
    if (!gen.empty()) {
        std::vector<uint32_t> bins(std::max<uint64_t>(nv, 1));
        if (bin_kind == "kmeans") {
            // k-means view binning (kmeans/kmeans.cpp), one view per cluster
            Timer kt;
            kt.start();
            uint32_t ncl = 0;
            if (g2h_kmeans(rp, nv, g2h_graph_cols(g), (uint32_t)nb, 1, 20, 0x6732A1D0C000000Bull, bins.data(), &ncl) != 0)
                die("kmeans: " + std::string(g2h_last_error()));
            kt.stop();
            say(kt.line("kmeans view binning, " + std::to_string(ncl) + " clusters"));
        } else if (g2h_kcore_bins(shells.data(), shells.size() - 1, (uint32_t)nb, bins.data(), nv) != 0) {
            die("kcore bins: " + std::string(g2h_last_error()));
        }
        pool.resize(nb);
        for (int b = 0; b < nb; b++) pool[b].id = b;
        for (uint64_t v = 0; v < nv; v++) pool[bins[v]].sources.push_back((uint32_t)v);
        for (int b = 0; b < nb; b++) {
            say("Source " + std::to_string(b) + ": " + std::to_string(pool[b].sources.size()));
            max_sources = std::max(max_sources, pool[b].sources.size());
        }
    } else {
        pool = load_bins(data + "." + std::to_string(hop) + ".hop.vbmap", nv, &max_sources);
    }
    if (ss)
        for (auto& b : pool) std::sort(b.sources.begin(), b.sources.end());
    if (dr) reorder_bins(pool, rp);
    if (ds) split_bins(pool, rp, cn, dst);
    for (size_t i = 0; i < pool.size(); i++) say(std::to_string(i) + " " + std::to_string(pool[i].id));

    Pipeline pipe;
    pipe.app = app;
    pipe.g = g;
    pipe.pool = &pool;
    pipe.hop = hop;
    pipe.root_degree = root_degree;
    pipe.pn = pn;
    pipe.cn = cn;
    pipe.qs = qs;
    pipe.gpus = gpus;
    // holders: a view never exceeds the whole graph
    pipe.cap = std::min<uint64_t>((uint64_t)(mem * 1e9) / 4, std::max<uint64_t>(ne, 1));
    pipe.plan = &plan;
    // memory budget per consumer: the device's memory shared by its consumers
    pipe.budget.assign(cn, 0);
    for (int c = 0; c < cn; c++) {
        uint64_t tot = 0;
        g2_device_info(c % gpus, nullptr, &tot, nullptr, 0);
        const int sharing = (cn - c % gpus + gpus - 1) / gpus;
        pipe.budget[c] = budget_gb > 0 ? (uint64_t)(budget_gb * 1e9) : (uint64_t)(tot * 0.9) / std::max(1, sharing);
    }
    std::vector<uint32_t> core;
    std::vector<std::pair<std::string, PipelineResult>> forms;
    if (sched == "plain" || sched == "both") {
        // the reference's schedule: source order, one root per wave, no donation
        pipe.form = BkForm{};
        if (app == App::BK) {
            pipe.form.tune.mce_donate_every = 0xffffffffu;
            pipe.form.tune.mce_roots_per_wave = 1;
        }
        if (want_learned) say("== plain schedule ==");
        forms.push_back({"plain", pipe.run()});
    }
    if (sched == "reverse") {
        // the reference's schedule with the sources backwards: one root per
        // wave, no donation, no cost model
        pipe.form = BkForm{};
        pipe.form.reverse = true;
        pipe.form.tune.mce_donate_every = 0xffffffffu;
        pipe.form.tune.mce_roots_per_wave = 1;
        forms.push_back({"reverse", pipe.run()});
    }
    if (want_learned) {
        say("== learned schedule ==");
        core.resize(nv);
        if (g2h_core_numbers(rp, nv, g2h_graph_cols(g), core.data()) != 0)
            die("core numbers: " + std::string(g2h_last_error()));
        pipe.form = BkForm{};
        pipe.form.learned = true;
        pipe.form.model = &model;
        pipe.form.core = core.data();
        pipe.form.nv = nv;
        pipe.form.tune.mce_donate_every = (uint32_t)donate;
        pipe.form.tune.mce_roots_per_wave = (uint32_t)rpw;
        pipe.form.tune.waves_per_cu = (uint32_t)wpc;
        forms.push_back({"learned", pipe.run()});
    }
    for (auto& f : forms) {
        if (f.second.total.count != forms[0].second.total.count || f.second.total.hist != forms[0].second.total.hist)
            die("the " + f.first + " schedule changed the result");
        // the two matching paths split a total over views differently
        if (app == App::GM && f.second.total.local_views && f.second.total.local_views != f.second.total.launches)
            die("views were counted by different matching paths (local bitmaps in " +
                std::to_string(f.second.total.local_views) + " of " + std::to_string(f.second.total.launches) +
                "); their per-view counts do not add up to the total");
    }
    const RunResult& res = forms[0].second.total;
    say("Total counts: " + std::to_string(res.count));
    if (app == App::BK) {
        std::string s = "Clique size histogram:";
        for (size_t k = 0; k < res.hist.size(); k++)
            if (res.hist[k]) s += " " + std::to_string(k) + ":" + std::to_string(res.hist[k]);
        say(s);
    }
    for (auto& f : forms) {
        std::ostringstream o;
        o << f.first << " schedule: pipeline " << f.second.seconds << " s, views running "
          << busy_seconds(f.second.total.spans) << " s, kernels " << f.second.total.kernel_ms << " ms";
        say(o.str());
    }
    if (forms.size() == 2) {
        std::ostringstream o;
        o << "learned vs plain: " << busy_seconds(forms[0].second.total.spans) / std::max(busy_seconds(forms[1].second.total.spans), 1e-9)
          << "x view run time, " << forms[0].second.seconds / std::max(forms[1].second.seconds, 1e-9)
          << "x pipeline time, " << forms[0].second.total.kernel_ms / std::max(forms[1].second.total.kernel_ms, 1e-9)
          << "x kernel time";
        say(o.str());
    }
    if (!json.empty()) {
        std::ofstream f(json);
        f << "{\"app\": \"" << (app == App::BK ? "bk" : "run") << "\", \"graph\": \"" << (gen.empty() ? data : gen)
          << "\", \"query\": " << query << ", \"vertices\": " << nv << ", \"edges\": " << ne / 2
          << ", \"total\": " << res.count << ", \"seconds\": " << forms[0].second.seconds << ", \"hist\": ";
        json_hist(f, res.hist);
        for (auto& fm : forms) {
            f << ", \"" << fm.first << "\": ";
            json_form(f, fm.second);
        }
        f << "}\n";
    }
    g2h_graph_free(g);
    return 0;
}

}  // namespace g2w
