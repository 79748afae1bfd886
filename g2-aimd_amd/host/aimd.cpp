// host/aimd.cpp -- the additive-increase / multiplicative-decrease controller
// of the root chunk, restated from system/subgraph_container.h:178-211 with
// the constants of common/gpu_env.h:42-69.
#include <algorithm>

#include "common.hpp"

namespace {
constexpr uint32_t kMaxChunk = 1000000, kMinChunk = 1, kDecFactor = 8, kMinStep = 1, kMinThreshold = 2000;
constexpr double kAddPc = 0.05;
}  // namespace

extern "C" {

void g2h_aimd_init(g2h_aimd* a) {
    a->chunk = kMaxChunk;
    a->step = kMinStep;
    a->add_step = kMinStep;
    a->threshold = kMinThreshold;
    a->recover = 1;
}

const char* g2h_aimd_adjust(g2h_aimd* a, int overflow) {
    const char* tag;
    if (overflow && a->chunk > kMinChunk) {
        a->threshold = std::max(kMinThreshold, a->chunk / 2);
        a->chunk = std::max(a->chunk / kDecFactor, kMinChunk);
        a->step = std::max((uint32_t)(a->chunk * kAddPc), kMinStep);
        a->add_step = a->step;
        tag = "--- chunk: ";
    } else if (!overflow && a->recover && a->chunk < kMaxChunk) {
        if (a->chunk < a->threshold) a->chunk = std::min(a->chunk + a->step, kMaxChunk);
        else a->chunk = std::min(a->chunk + a->add_step, kMaxChunk);
        a->step *= 2;
        tag = "+++ chunk: ";
    } else {
        tag = "... chunk: ";
    }
    a->recover = 1;
    return tag;
}

}  // extern "C"
