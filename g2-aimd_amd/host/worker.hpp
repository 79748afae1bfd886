// host/worker.hpp -- the reference's worker pipeline over the two C ABIs.
//
// Same command line, same producer/consumer structure and the same output
// lines as Worker<App>::run (system/worker.h:28-280):
//   * CommandLine flags -dg -q -m -t -qs -h -pn -cn -dr -ds -dst -ss
//     (system/worker.h:31-43, common/command_line.h);
//   * the .bin graph, the <graph>.<hop>.hop.vbmap bin map, sort / reorder /
//     split of the bins (view/view_bin_manager.h:34-189);
//   * pn producer threads materialise bins into qs pinned holders
//     (view/view_bin.h:61-159), cn consumer threads pull holders over two
//     bounded queues (released / assigned) and run the app on the device.
// Differences by design:
//   * consumers are not limited to one per GPU: consumer c uses device
//     c % #GPUs with an explicit memory budget (the reference grabs all free
//     memory per executor, system/appbase.h:93-114);
//   * `bk` prints the summed maximal-clique count as "Total counts" (the
//     reference's device code never writes the counters that line sums, so it
//     prints 0 there) and the clique-size histogram;
//   * a view's roots run in one launch; the reference's AIMD chunk loop
//     (system/pipeline_executor.h:151-207) starts only after a launch does
//     not fit the memory budget.
// Additions (no reference counterpart):
//   -gen social:<n>[:seed] | rmat:<scale>[:seed] | chunglu:<n>[:seed]
//        a seeded synthetic graph instead of -dg, degeneracy-recoded and cut
//        into -nb (default 4) bins in memory (no files): -bins kcore (the
//        kcore tool's blocks) or -bins kmeans (k-means view binning, 20
//        rounds, fixed seed)
//   -sched plain | learned | both     (bk) root schedule of every view:
//        plain = source order, one root per wave, no work donation;
//        learned = cost-model order and heavy-root splits, work donation,
//        small roots batched several per wave (-rpw), -wpc resident waves
//        per CU; both = the pipeline twice, plain then learned
//   -model <file>   trained cost model (standardiser + MLP parameters)
//   -train <file>   train the cost model first, on plain runs of the
//        -train-gen graphs (comma-separated specs), and write it to <file>
//   -epochs <n>     training epochs (20)
//   -qf <file>      (run) a pattern file, one edge "a b" per line, instead of
//        a -q preset; its symmetry-breaking conditions come from the
//        pattern's own automorphism group
//   -budget <GB>    device memory per consumer
//   -json <file>    totals, histogram and per-form timings as JSON
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <iostream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/time.h>
#include <thread>
#include <vector>

#include "g2hip.h"
#include "g2host.h"

namespace g2w {

// ---- command line (common/command_line.h semantics) ----------------------
struct CommandLine {
    int argc;
    char** argv;
    std::string get(const std::string& opt, const std::string& def) const {
        for (int i = 1; i + 1 < argc; i++)
            if (opt == argv[i]) return argv[i + 1];
        return def;
    }
    int get_int(const std::string& opt, int def) const {
        for (int i = 1; i + 1 < argc; i++)
            if (opt == argv[i]) return std::atoi(argv[i + 1]);
        return def;
    }
    double get_double(const std::string& opt, double def) const {
        for (int i = 1; i + 1 < argc; i++)
            if (opt == argv[i]) return std::atof(argv[i + 1]);
        return def;
    }
};

// ---- "finish <tag>, elapsed_time=<ms>ms" (common/timer.h) -------------------
struct Timer {
    double t0 = 0, t1 = 0;
    static double now_us() {
        timeval tv;
        gettimeofday(&tv, nullptr);
        return tv.tv_sec * 1e6 + tv.tv_usec;
    }
    void start() { t0 = now_us(); }
    void stop() { t1 = now_us(); }
    double ms() const { return (t1 - t0) / 1000.0; }
    std::string line(const std::string& tag) const {
        std::ostringstream o;
        o << std::fixed << "finish " << tag << ", elapsed_time=" << ms() << "ms";
        return o.str();
    }
};

// whole lines go to stdout atomically across threads
inline void say(const std::string& s) {
    static std::mutex m;
    std::lock_guard<std::mutex> g(m);
    std::cout << s << std::endl;
}

[[noreturn]] inline void die(const std::string& s) {
    say(s);
    std::exit(-1);
}

// ---- bounded blocking queue (the MPMC hand-off of holder ids) --------------
class BoundedQueue {
   public:
    explicit BoundedQueue(size_t cap) : cap_(cap) {}
    void push(int v) {
        std::unique_lock<std::mutex> l(m_);
        not_full_.wait(l, [&] { return q_.size() < cap_; });
        q_.push_back(v);
        not_empty_.notify_one();
    }
    void pop(int& v) {
        std::unique_lock<std::mutex> l(m_);
        not_empty_.wait(l, [&] { return !q_.empty(); });
        v = q_.front();
        q_.pop_front();
        not_full_.notify_one();
    }
    bool try_pop(int& v) {
        std::lock_guard<std::mutex> l(m_);
        if (q_.empty()) return false;
        v = q_.front();
        q_.pop_front();
        not_full_.notify_one();
        return true;
    }

   private:
    size_t cap_;
    std::deque<int> q_;
    std::mutex m_;
    std::condition_variable not_full_, not_empty_;
};

// ---- view bins -------------------------------------------------------------
struct ViewBin {
    int id;
    std::vector<uint32_t> sources;
    uint64_t factor(const uint64_t* rp) const {  // sum of source degrees
        uint64_t s = 0;
        for (uint32_t v : sources) s += rp[v + 1] - rp[v];
        return s;
    }
};

struct Holder {  // pinned materialised view (view/view_bin_holder.h)
    uint64_t nv = 0, cap = 0, ne = 0;
    uint64_t* rp = nullptr;
    uint32_t* cols = nullptr;
    int view_id = -1;
    std::vector<uint32_t> sources;
    Holder(uint64_t nv_, uint64_t cap_) : nv(nv_), cap(cap_) {
        rp = (uint64_t*)g2_pinned_alloc((nv + 1) * 8);
        cols = (uint32_t*)g2_pinned_alloc(std::max<uint64_t>(cap, 1) * 4);
        if (!rp || !cols) die("cudaMallocHost failed: pinned view holder");
    }
    ~Holder() {
        g2_pinned_free(rp);
        g2_pinned_free(cols);
    }
};

std::vector<ViewBin> load_bins(const std::string& file, uint64_t nv, size_t* max_sources);
void split_bins(std::vector<ViewBin>& pool, const uint64_t* rp, int device_num, int times);
void reorder_bins(std::vector<ViewBin>& pool, const uint64_t* rp);

// ---- per-consumer device executor (system/pipeline_executor.h) -------------
struct RunResult {
    uint64_t count = 0;                 // GM matches / BK maximal cliques of this view
    std::vector<uint64_t> hist;         // BK clique sizes
    double kernel_ms = 0;               // HIP-event time of the search kernels
    double sched_ms = 0;                // learned schedule: features, inference, order
    double idle_sum = 0;                // sum of per-launch wave idle fractions
    uint32_t launches = 0;              // g2_mce_run / g2_gm_run calls that counted
    uint32_t retries = 0;               // launches refused by the memory budget
    uint64_t nodes = 0, donations = 0, split_roots = 0;
    uint64_t host_spill_bytes = 0;      // largest host spill of a launch
    uint32_t local_views = 0;           // GM: views counted by the local-bitmap path
    uint64_t device_bytes = 0;          // largest device memory held after a launch
    std::vector<std::pair<double, double>> spans;  // wall-clock (us) of each view's run, schedule included
    void add(const RunResult& r);
};

// cost model shared by the consumers of a learned run
struct Model {
    g2h_norm norm{};
    std::vector<float> params;
};

// how a consumer schedules the roots of its views
struct BkForm {
    bool learned = false;
    bool reverse = false;               // sources in reverse order, no cost model
    const Model* model = nullptr;       // learned form only
    const uint32_t* core = nullptr;     // core numbers of the whole graph (learned form)
    uint64_t nv = 0;
    g2_tuning tune{};
};

class Executor {
   public:
    Executor(int consumer, int device, uint64_t budget, const g2_tuning& tune = g2_tuning{});
    ~Executor();
    void transfer(const Holder& h);
    RunResult run_bk(uint32_t hist_len, const BkForm& form);
    RunResult run_gm();
    void set_plan(const g2_plan& p);

   private:
    void check(int rc, const char* what);
    int consumer_;
    int device_;
    std::string tag_;
    g2_ctx* ctx_ = nullptr;
    g2_mlp* mlp_ = nullptr;
    uint64_t ns_ = 0;
    std::vector<uint32_t> sources_;
};

// synthetic input of -gen / -train-gen: preprocessed, degeneracy-recoded,
// shell offsets for the kcore bins
struct GenGraph {
    g2h_graph* g = nullptr;
    std::vector<uint64_t> shells;
};
GenGraph generate(const std::string& spec);
// train the cost model on plain runs of the given graphs (specs) and return it
Model train_model(const std::vector<std::string>& specs, int device, uint32_t epochs);

// ---- the worker ------------------------------------------------------------
enum class App { BK, GM };
int run_worker(int argc, char** argv, App app);

}  // namespace g2w
