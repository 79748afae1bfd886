// csrc/g2_noalloc.hip -- allocation-free matching entry point.
//
// g2_gm_run_noalloc counts a plan over caller-owned device buffers without
// allocating device memory: its counters and per-wave candidate-list arenas
// live in `scratch`, a file-scope array each device holds until the process
// ends.  Every call on a device uses that device's copy, so calls that overlap
// on one device race.
#include <algorithm>
#include <cstdio>
#include <mutex>

#include "g2_device.hpp"
#include "g2_internal.hpp"

namespace g2 {
constexpr uint64_t kNoAllocBytes = 64ull << 20;  // the array's size: 64 MiB
constexpr uint64_t kNoAllocHead = 256;           // counters at the front; the arenas follow
constexpr int kNoAllocMaxDevices = 64;
// counters after GCTR_*: largest degree, malformed row_ptrs / column / source
enum { NCHK_MAXDEG = GCTR_N, NCHK_BADROW, NCHK_BADCOL, NCHK_BADSRC, NCHK_N };
static_assert(NCHK_N * 8 <= kNoAllocHead, "counters fit the head");
}  // namespace g2

static __device__ __align__(256) uint64_t scratch[g2::kNoAllocBytes / 8];

namespace g2 {

// the counters must start at zero on every call (a kernel, not a memset, so
// that the sequence stays capturable)
__global__ __launch_bounds__(64) void noalloc_clear_kernel(unsigned long long* head) {
    if (threadIdx.x < kNoAllocHead / 8) head[threadIdx.x] = 0;
}

// Largest degree and the bounds of the caller's CSR, read only within the
// sizes the caller gave: the search kernel indexes without checks, so it
// runs only when these come back clean.
__global__ __launch_bounds__(256) void noalloc_check_kernel(const uint64_t* __restrict__ rp, uint64_t nv,
                                                            const uint32_t* __restrict__ cols, uint64_t ne,
                                                            const uint32_t* __restrict__ src, uint64_t ns,
                                                            unsigned long long* __restrict__ ctr) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t md = 0, bad_row = 0, bad_col = 0, bad_src = 0;
    if (tid == 0 && rp[nv] != ne) bad_row = 1;
    for (uint64_t v = tid; v < nv; v += stride) {
        const uint64_t a = rp[v], b = rp[v + 1];
        if (b < a || b > ne) bad_row = 1;
        else md = b - a > md ? b - a : md;
    }
    for (uint64_t j = tid; j < ne; j += stride) bad_col |= cols[j] >= nv;
    for (uint64_t i = tid; i < ns; i += stride) bad_src |= src[i] >= nv;
    md = wave_max(md);
    bad_row = wave_or(bad_row);
    bad_col = wave_or(bad_col);
    bad_src = wave_or(bad_src);
    if (lane_id() == 0) {
        if (md) atomicMax(&ctr[NCHK_MAXDEG], (unsigned long long)md);
        if (bad_row) atomicOr(&ctr[NCHK_BADROW], 1ull);
        if (bad_col) atomicOr(&ctr[NCHK_BADCOL], 1ull);
        if (bad_src) atomicOr(&ctr[NCHK_BADSRC], 1ull);
    }
}

struct NoAllocDev {
    uint64_t* ptr = nullptr;   // device address of `scratch` on this device
    uint32_t num_cus = 0;
};
static NoAllocDev g_noalloc[kNoAllocMaxDevices];
static std::mutex g_noalloc_mu;
static thread_local char g_noalloc_err[160];

static int noalloc_fail(int code, const char* what, hipError_t e = hipSuccess) {
    if (e != hipSuccess) std::snprintf(g_noalloc_err, sizeof(g_noalloc_err), "%s: %s", what, hipGetErrorString(e));
    else std::snprintf(g_noalloc_err, sizeof(g_noalloc_err), "%s", what);
    return code;
}

// Once per device, with that device current: the device address of its copy
// of `scratch` (the array's name in host code is only a stand-in) and its CU
// count.
static int noalloc_device(int device, NoAllocDev* out) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDev& d = g_noalloc[device];
    if (!d.ptr) {
        hipError_t e = hipGetSymbolAddress((void**)&d.ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            d.ptr = nullptr;
            return noalloc_fail(G2_EHIP, "hipGetSymbolAddress(scratch)", e);
        }
        int cus = 0;
        if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) {
            d.ptr = nullptr;
            return noalloc_fail(G2_EHIP, "CU count", e);
        }
        d.num_cus = (uint32_t)cus;
    }
    *out = d;
    return G2_OK;
}

}  // namespace g2

extern "C" {

int g2_gm_run_noalloc(int device, const uint64_t* row_ptrs, uint64_t nv, const uint32_t* cols, uint64_t ne,
                      const uint32_t* sources, uint64_t ns, const g2_plan* plan, uint64_t* count) {
    using namespace g2;
    if (!row_ptrs || (ne && !cols) || (ns && !sources) || !plan || !count)
        return noalloc_fail(G2_EINVAL, "null argument");
    if (nv >= 0xffffffffull) return noalloc_fail(G2_EINVAL, "vertex ids must fit in u32");
    if (device < 0 || device >= kNoAllocMaxDevices) return noalloc_fail(G2_EINVAL, "device out of range");
    GmProgram prog{};
    const char* why = "";
    if (gm_program(plan, &prog, &why) != G2_OK) return noalloc_fail(G2_EINVAL, why);
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return noalloc_fail(G2_EHIP, "hipSetDevice", e);
    NoAllocDev d;
    int rc = noalloc_device(device, &d);
    if (rc) return rc;
    auto* ctr = (unsigned long long*)d.ptr;
    hipLaunchKernelGGL(noalloc_clear_kernel, dim3(1), dim3(64), 0, 0, ctr);
    const uint64_t n_max = std::max(std::max(nv, ne), ns);
    const uint64_t cgrid = std::max<uint64_t>(1, std::min<uint64_t>((n_max + 255) / 256, 4096));
    hipLaunchKernelGGL(noalloc_check_kernel, dim3(cgrid), dim3(256), 0, 0, row_ptrs, nv, cols, ne, sources, ns, ctr);
    if ((e = hipGetLastError()) != hipSuccess) return noalloc_fail(G2_EHIP, "check launch", e);
    unsigned long long h[NCHK_N];
    if ((e = hipMemcpy(h, ctr, sizeof(h), hipMemcpyDeviceToHost)) != hipSuccess)
        return noalloc_fail(G2_EHIP, "D2H check", e);
    if (h[NCHK_BADROW]) return noalloc_fail(G2_EINVAL, "row_ptrs must be non-decreasing, within [0, ne], and end at ne");
    if (h[NCHK_BADCOL]) return noalloc_fail(G2_EINVAL, "column id out of range");
    if (h[NCHK_BADSRC]) return noalloc_fail(G2_EINVAL, "source id out of range");
    // one candidate list and one pre-intersection per position per wave, each
    // as long as the longest row, as in g2_gm_run
    const uint64_t cap = std::max<uint64_t>(h[NCHK_MAXDEG], 64);
    const uint64_t per_wave = cap * 2 * prog.n * 4;
    const uint64_t fit = (kNoAllocBytes - kNoAllocHead) / per_wave;
    if (fit == 0) return noalloc_fail(G2_EBUDGET, "the largest degree needs a wave arena beyond the 64 MiB scratch");
    const uint64_t waves = std::min<uint64_t>((uint64_t)d.num_cus * WK_GM, fit);
    const uint32_t bw = (uint32_t)std::min<uint64_t>(kGmMaxBlockWaves, waves);
    const uint64_t blocks = waves / bw;
    GmArgs a{};
    a.rp = row_ptrs;
    a.cols = cols;
    a.src = sources;
    a.task_off = nullptr;
    a.ns = ns;
    a.ntasks = ns;
    a.chunk = 0;
    a.scratch = (uint32_t*)((char*)d.ptr + kNoAllocHead);
    a.cap = cap;
    a.ctr = ctr;
    a.per_root = nullptr;
    if ((e = gm_launch(a, prog, blocks, bw, 0)) != hipSuccess) return noalloc_fail(G2_EHIP, "gm launch", e);
    if ((e = hipMemcpy(h, ctr, GCTR_N * 8, hipMemcpyDeviceToHost)) != hipSuccess)
        return noalloc_fail(G2_EHIP, "D2H count", e);
    *count = h[GCTR_COUNT];
    return G2_OK;
}

const char* g2_noalloc_last_error(void) { return g2::g_noalloc_err; }

}  // extern "C"
