// csrc/g2_learn.hip -- root features and the cost model of the learned root
// scheduler (no counterpart in the reference, whose roots run in source order).
//
// * feature_kernel: one wave64 per root computes integer features
//     deg(v), the distinct 2-hop ball size |{v} u N(v) u N(N(v))| (a
//     HyperLogLog estimate from the registers sketch_kernel leaves per
//     vertex), triangles t(v), neighbours sharing >= 1 common neighbour with
//     v, and p0 = |{u > v}| (the size of the root's P set in the
//     maximal-clique search tree);
//   triangles and triangle-carrying neighbours come from one pass over the
//   view's oriented rows (local_triangle_features: each triangle found once,
//   at its lowest (degree, id) rank) when every source's neighbours have
//   their rows; otherwise each root streams the shorter of N(v), N(u) and
//   binary searches the other, one neighbour u at a time.
// * core_step_kernel: core numbers by the h-index iteration;
//   inputs_kernel: the cost model's 8 inputs in fp32.
// * mlp: 8 -> 64 -> 64 -> 1 with ReLU, BF16 weights and activations on the
//   matrix cores (v_mfma_f32_16x16x16bf16_1k), FP32 accumulation, FP32 master
//   weights and Adam.  A wave owns 16 samples: layer 1 is 4 MFMA tiles (K padded
//   8 -> 16), layer 2 is 4x4 tiles, layer 3 a 16-lane reduction.  The backward
//   pass forms dW2 = h1^T dh2, dW1 = x^T dh1 and dh1 = dh2 W2^T on the same
//   instruction and accumulates each wave's gradient in registers before one
//   atomic flush.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "g2_device.hpp"
#include "g2_internal.hpp"

namespace g2 {

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
struct FeatArgs {
    const uint64_t* __restrict__ rp;
    const uint32_t* __restrict__ cols;
    const uint32_t* __restrict__ src;
    uint64_t ns;
    unsigned long long* next;
    int64_t* out;  // ns x G2_NFEAT_INT
    // triangles and triangle-carrying neighbours per rank (null: counted here)
    const unsigned long long* tri;
    const uint32_t* sup;
    const uint32_t* rank;
    const uint8_t* sketch;  // nv x 64 HyperLogLog registers of the closed rows N[u]
    double lc[65];          // linear counting: lc[V] = 64 ln(64 / V), from the host's libm
};

// HyperLogLog of vertex sets (Flajolet, Fusy, Gandouet, Meunier 2007) with
// m = 64 registers: hash = the splitmix64 finaliser of the id, register =
// its top 6 bits, rank = leading zeros of the other 58 bits + 1, capped at 32
// so that sum 2^(32 - reg) is an exact integer.
__device__ inline uint64_t hll_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ inline void hll_put(uint32_t* reg, uint32_t w) {
    const uint64_t h = hll_mix(w);
    const uint64_t rest = h << 6;
    const uint32_t rho = rest ? umin32((uint32_t)__builtin_clzll(rest) + 1, 32u) : 32u;
    atomicMax(&reg[h >> 58], rho);
}

// registers of every vertex's closed row N[u] = N(u) u {u}: one wave per
// vertex, the maxima taken in the wave's 64 LDS words
__global__ __launch_bounds__(256) void sketch_kernel(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols,
                                                     uint64_t nv, uint8_t* __restrict__ sk) {
    __shared__ uint32_t regs[4][kWave];
    const uint32_t lane = lane_id();
    uint32_t* reg = regs[threadIdx.x / kWave];
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / kWave);
    for (uint64_t u = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < nv; u += nw) {
        reg[lane] = 0;
        wave_sync();
        if (lane == 0) hll_put(reg, (uint32_t)u);
        const uint64_t lo = rp[u], hi = rp[u + 1];
        for (uint64_t e = lo + lane; e < hi; e += kWave) hll_put(reg, cols[e]);
        wave_sync();
        sk[u * kWave + lane] = (uint8_t)reg[lane];
        wave_sync();
    }
}

__global__ __launch_bounds__(256) void feature_kernel(FeatArgs a) {
    const uint32_t lane = lane_id();
    // sources come in runs of 16 per counter fetch: a root's features cost a
    // few row walks, and one atomic per root would queue the waves on it
    constexpr uint32_t kRun = 16;
    uint64_t t = 0, tend = 0;
    while (true) {
        if (t == tend) {
            t = fetch_tasks(a.next, kRun);
            tend = umin64(t + kRun, a.ns);
        }
        if (t >= a.ns) break;
        const uint64_t ti = t++;
        const uint32_t v = uniform(a.src[ti]);
        const uint64_t lo = a.rp[v], hi = a.rp[v + 1], d = hi - lo;
        uint64_t p0 = 0, tri2 = 0, shared = 0;
        for (uint64_t e = lo + lane; e < hi; e += kWave) p0 += a.cols[e] > v;
        // 2-hop ball: register-wise maximum over N[v] and the N[u], u in N(v)
        // (lane l holds register l; 8 neighbours' registers in flight)
        uint32_t reg = a.sketch[(uint64_t)v * kWave + lane];
        for (uint64_t e0 = lo; e0 < hi; e0 += 8) {
            uint32_t r[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint64_t e = e0 + q;
                r[q] = e < hi ? a.sketch[(uint64_t)uniform(a.cols[e]) * kWave + lane] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) reg = umax32(reg, r[q]);
        }
        const uint64_t hsum = wave_sum(1ull << (32 - reg));
        const uint64_t zeros = wave_sum(reg == 0);
        for (uint64_t e = a.tri ? hi : lo; e < hi; e++) {
            const uint32_t u = a.cols[e];
            const uint64_t ulo = a.rp[u], uhi = a.rp[u + 1];
            uint64_t c = 0;
            if (uhi - ulo <= d) {
                for (uint64_t k = ulo + lane; k < uhi; k += kWave) c += contains(a.cols, lo, hi, a.cols[k]);
            } else {
                for (uint64_t k = lo + lane; k < hi; k += kWave) c += contains(a.cols, ulo, uhi, a.cols[k]);
            }
            c = wave_sum(c);
            tri2 += c;
            shared += c > 0;
        }
        p0 = wave_sum(p0);
        if (lane == 0) {
            // E = alpha_64 m^2 / sum 2^-reg; linear counting below 5m/2 when a
            // register is empty
            double est = (0.709 * 4096.0 * 4294967296.0) / (double)hsum;
            if (est <= 160.0 && zeros > 0) est = a.lc[zeros];
            int64_t* o = a.out + ti * G2_NFEAT_INT;
            o[0] = (int64_t)d;
            o[1] = (int64_t)(est + 0.5);
            o[2] = a.tri ? (int64_t)a.tri[a.rank[v]] : (int64_t)(tri2 / 2);
            o[3] = a.tri ? (int64_t)a.sup[a.rank[v]] : (int64_t)shared;
            o[4] = (int64_t)p0;
        }
    }
}

// ---------------------------------------------------------------------------
// core numbers: the h-index iteration (Lu, Zhou, Zhang, Stanley, "The H-index
// of a network node and its relation to degree and coreness", Nature
// Communications 7, 2016): starting from the degrees, each vertex's value
// becomes the h-index of its neighbours' values (the largest k with at least
// k neighbours valued >= k) until nothing changes; the values only decrease,
// stay >= the core number, and end at it.  Updates are applied in place
// (asynchronously); a vertex is revisited only when a neighbour decreased.
// ---------------------------------------------------------------------------
__device__ inline uint32_t count_at_least(const uint32_t* __restrict__ cols, uint64_t lo, uint64_t hi,
                                          const uint32_t* h, uint32_t k) {
    uint32_t c = 0;
    for (uint64_t e = lo + lane_id(); e < hi; e += kWave) c += __hip_atomic_load(&h[cols[e]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= k;
    return (uint32_t)wave_sum(c);
}

__global__ __launch_bounds__(256) void core_init_kernel(const uint64_t* __restrict__ rp, uint64_t nv, uint32_t* h,
                                                        uint8_t* act) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        h[v] = (uint32_t)umin64(rp[v + 1] - rp[v], 0xffffffffull);
        act[v] = 1;
    }
}

__global__ __launch_bounds__(256) void core_clear_kernel(uint8_t* act, uint64_t nv) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x)
        act[v] = 0;
}

__global__ __launch_bounds__(256) void core_step_kernel(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols,
                                                        uint64_t nv, uint32_t* h, const uint8_t* __restrict__ act_in,
                                                        uint8_t* act_out, unsigned long long* changed) {
    const uint64_t nw = (uint64_t)gridDim.x * blockDim.x / kWave;
    for (uint64_t v = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; v < nv; v += nw) {
        if (!act_in[v]) continue;
        const uint32_t cur = __hip_atomic_load(&h[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == 0) continue;
        const uint64_t lo = rp[v], hi = rp[v + 1];
        if (count_at_least(cols, lo, hi, h, cur) >= cur) continue;
        uint32_t a = 0, b = cur - 1;  // k = 0 always holds, k = cur does not
        while (a < b) {
            const uint32_t mid = a + (b - a + 1) / 2;
            if (count_at_least(cols, lo, hi, h, mid) >= mid) a = mid;
            else b = mid - 1;
        }
        if (lane_id() == 0) {
            __hip_atomic_store(&h[v], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            atomicOr(changed, 1ull);
        }
        for (uint64_t e = lo + lane_id(); e < hi; e += kWave) act_out[cols[e]] = 1;
    }
}

// core numbers of the uploaded view's vertices into dcore (device, nv u32)
static int core_numbers_run(g2_ctx* c, uint32_t* dcore, uint8_t* act0, uint8_t* act1) {
    hipError_t e;
    const uint64_t nv = c->nv;
    const auto* rp = (const uint64_t*)c->rp.ptr;
    const auto* cols = (const uint32_t*)c->cols.ptr;
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((nv + 255) / 256, 8192));
    const uint64_t wgrid = std::max<uint64_t>(1, (uint64_t)c->num_cus * 32 / 4);
    auto* changed = (unsigned long long*)c->counters.ptr;
    hipLaunchKernelGGL(core_init_kernel, dim3(grid), dim3(256), 0, c->stream, rp, nv, dcore, act0);
    for (;;) {
        hipLaunchKernelGGL(core_clear_kernel, dim3(grid), dim3(256), 0, c->stream, act1, nv);
        if ((e = hipMemsetAsync(changed, 0, 8, c->stream))) return hip_err(c, e, "memset");
        hipLaunchKernelGGL(core_step_kernel, dim3(wgrid), dim3(256), 0, c->stream, rp, cols, nv, dcore, act0, act1, changed);
        if ((e = hipGetLastError())) return hip_err(c, e, "core launch");
        unsigned long long h = 0;
        if ((e = hipMemcpyAsync(&h, changed, 8, hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H core");
        if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "core sync");
        if (!h) break;
        std::swap(act0, act1);
    }
    return G2_OK;
}

// The cost model's 8 inputs in fp32, from the integer features, the core
// numbers and the source ids (the fp32 formula, one rounding per operation)
__global__ __launch_bounds__(256) void inputs_kernel(const int64_t* __restrict__ feat, const uint32_t* __restrict__ src,
                                                     const uint32_t* __restrict__ core, uint64_t ns, uint64_t nv,
                                                     float* __restrict__ x) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t* f = feat + i * G2_NFEAT_INT;
        const float d = (float)f[0], tri = (float)f[2];
        const float pairs = (float)(f[0] * (f[0] - 1));
        float* o = x + i * 8;
        o[0] = log1pf(d);
        o[1] = log1pf((float)core[src[i]]);
        o[2] = log1pf((float)f[1]);
        o[3] = log1pf(tri);
        o[4] = f[0] > 1 ? (2.0f * tri) / pairs : 0.0f;
        o[5] = (float)src[i] / (float)(nv ? nv : 1);
        o[6] = (float)f[3] / fmaxf(d, 1.0f);
        o[7] = log1pf((float)f[4]);
    }
}

// ---- the learned schedule on the device ------------------------------------
// standardised inputs, as g2h_norm_apply: (float)(((double)x - mu) / sd)
__global__ __launch_bounds__(256) void norm_kernel(float* x, uint64_t n, const float* __restrict__ musd) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * 8; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)(i & 7);
        x[i] = (float)(((double)x[i] - (double)musd[k]) / (double)musd[8 + k]);
    }
}

// clamp(pred, 0, 60) (NaN: 0), the predicted cost of a root
__device__ inline float sched_p(float y) { return isfinite(y) ? fminf(fmaxf(y, 0.0f), 60.0f) : 0.0f; }

// order keys ~bits(p) (ascending = longest predicted first) with the source
// index as value, and per-block sums of the costs expm1(p) in a fixed order
__global__ __launch_bounds__(256) void sched_keys_kernel(const float* __restrict__ pred, uint64_t n, uint32_t* keys,
                                                         uint32_t* idx, double* partial) {
    __shared__ double red[256];
    double acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float p = sched_p(pred[i]);
        keys[i] = ~__float_as_uint(p);
        idx[i] = (uint32_t)i;
        acc += expm1((double)p);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// parts per root: ceil(cost / share), at most its later neighbours (the
// first-level children it can have) and max_split, at least 1
__global__ __launch_bounds__(256) void sched_split_kernel(const float* __restrict__ pred, const int64_t* __restrict__ feat,
                                                          uint64_t n, double share, uint32_t max_split, uint32_t* split) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t later = feat[i * G2_NFEAT_INT + 4];
        double k = ceil(expm1((double)sched_p(pred[i])) / share);
        k = fmin(k, (double)(later > 1 ? later : 1));
        k = fmin(fmax(k, 1.0), (double)max_split);
        split[i] = (uint32_t)k;
    }
}

// features of every uploaded source into dfeat (device, ns x G2_NFEAT_INT);
// returns after the device has finished (its triangle buffers are freed)
static int features_dev(g2_ctx* c, int64_t* dfeat) {
    int rc;
    if ((rc = ensure(c, c->counters, 8, "counters"))) return rc;
    hipError_t e;
    // triangles through the oriented rows when every source's neighbours have
    // their rows (whole graph; views materialised with hop >= 2) and the
    // budget holds the rows; the rows are dropped again unless a matching
    // run had built them
    const unsigned long long* tri = nullptr;
    const uint32_t* sup = nullptr;
    const uint32_t* rank = nullptr;
    const bool had_orient = c->orient_valid;
    struct TriBuf {
        g2_ctx* c;
        DevBuf b;
        bool drop;
        ~TriBuf() {
            if (b.ptr) (void)hipFree(b.ptr);
            c->used -= b.bytes;
            if (drop) orient_release(c);
        }
    } tri_buf{c, {}, false};
    const uint64_t nv1 = std::max<uint64_t>(c->nv, 1), ne1 = std::max<uint64_t>(c->ne, 1);
    const bool closed = c->sources_all || c->unclosed_max_degree == 0;
    if (closed && c->ns && c->nv + 1 <= 0x7fffffffull && c->ne < (1ull << 32)) {
        auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
        const uint64_t o_sup = al(nv1 * 8), o_flag = o_sup + al(nv1 * 4);
        if (ensure(c, tri_buf.b, o_flag + ne1, "triangle features") == G2_OK) {
            char* base = (char*)tri_buf.b.ptr;
            tri_buf.drop = !had_orient;
            if (local_triangle_features(c, (unsigned long long*)base, (uint32_t*)(base + o_sup),
                                        (uint8_t*)(base + o_flag), ne1, &rank) == G2_OK) {
                tri = (const unsigned long long*)base;
                sup = (const uint32_t*)(base + o_sup);
            }
        }
    }
    // HyperLogLog registers of every vertex's closed row (nv x 64 B, for this call)
    struct SketchBuf {
        g2_ctx* c;
        DevBuf b;
        ~SketchBuf() {
            if (b.ptr) (void)hipFree(b.ptr);
            c->used -= b.bytes;
        }
    } sk_buf{c, {}};
    if ((rc = ensure(c, sk_buf.b, nv1 * kWave, "2-hop sketches"))) return rc;
    const uint64_t blocks = std::max<uint64_t>(1, (uint64_t)c->num_cus * waves_per_cu(c, WK_FEATURES) / 4);
    if (c->nv) hipLaunchKernelGGL(sketch_kernel, dim3(blocks), dim3(256), 0, c->stream, (const uint64_t*)c->rp.ptr,
                                  (const uint32_t*)c->cols.ptr, c->nv, (uint8_t*)sk_buf.b.ptr);
    if ((e = hipGetLastError())) return hip_err(c, e, "sketch launch");
    if ((e = hipMemsetAsync(c->counters.ptr, 0, 8, c->stream))) return hip_err(c, e, "memset");
    FeatArgs a{(const uint64_t*)c->rp.ptr, (const uint32_t*)c->cols.ptr, (const uint32_t*)c->src.ptr, c->ns,
               (unsigned long long*)c->counters.ptr, dfeat, tri, sup, rank, (const uint8_t*)sk_buf.b.ptr, {}};
    a.lc[0] = 0.0;
    for (int z = 1; z <= 64; z++) a.lc[z] = 64.0 * std::log(64.0 / (double)z);
    hipLaunchKernelGGL(feature_kernel, dim3(blocks), dim3(256), 0, c->stream, a);
    if ((e = hipGetLastError())) return hip_err(c, e, "feature launch");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "feature sync");
    return G2_OK;
}

int features_run(g2_ctx* c, int64_t* out) {
    // the output buffer lives for this call only, on every return path
    struct CallBuf {
        g2_ctx* c;
        DevBuf b;
        ~CallBuf() {
            if (b.ptr) (void)hipFree(b.ptr);
            c->used -= b.bytes;
        }
    } out_buf{c, {}};
    DevBuf& dout = out_buf.b;
    int rc;
    if ((rc = ensure(c, dout, std::max<uint64_t>(c->ns, 1) * G2_NFEAT_INT * 8, "features"))) return rc;
    if ((rc = features_dev(c, (int64_t*)dout.ptr))) return rc;
    hipError_t e;
    if (c->ns && (e = hipMemcpy(out, dout.ptr, c->ns * G2_NFEAT_INT * 8, hipMemcpyDeviceToHost)))
        return hip_err(c, e, "D2H features");
    return G2_OK;
}

// ---------------------------------------------------------------------------
// MLP 8 -> 64 -> 64 -> 1 on bf16 MFMA
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 8, kH = 64;
// parameter offsets (fp32 master copy)
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH, kNParam = oB3 + 1;
static_assert(kNParam == G2_MLP_NPARAM, "parameter count");

// float -> bf16, round to nearest even (NaN-free inputs)
__device__ __host__ inline uint16_t f2bf(float f) {
    uint32_t x;
    __builtin_memcpy(&x, &f, 4);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
__device__ __host__ inline float bf2f(uint16_t b) {
    uint32_t x = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &x, 4);
    return f;
}

struct MlpShared {
    uint16_t w1[16][kH];     // rows >= kIn are zero (K padded to 16)
    uint16_t w2[kH][kH];
    float w3[kH];            // bf16-rounded, kept as fp32 for the VALU dot
    float b1[kH], b2[kH], b3;
};

__device__ inline void load_weights(MlpShared& S, const float* __restrict__ p) {
    for (int i = threadIdx.x; i < 16 * kH; i += blockDim.x) {
        const int k = i / kH, n = i % kH;
        S.w1[k][n] = k < kIn ? f2bf(p[oW1 + k * kH + n]) : 0;
    }
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) S.w2[i / kH][i % kH] = f2bf(p[oW2 + i]);
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        S.w3[i] = bf2f(f2bf(p[oW3 + i]));
        S.b1[i] = p[oB1 + i];
        S.b2[i] = p[oB2 + i];
    }
    if (threadIdx.x == 0) S.b3 = p[oB3];
    __syncthreads();
}

struct WaveTile {  // per-wave LDS for one tile of 16 samples
    uint16_t x[16][16];      // bf16 inputs, K padded
    uint16_t h1[16][kH];     // bf16 post-ReLU layer-1 activations
    uint16_t g[16][kH];      // bf16 gradient staging (dh2, then dh1)
    float y[16];
};

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// forward of one 16-sample tile; leaves h1 (bf16) in T and h2 (fp32, D layout:
// lane holds rows 4*(lane/16)+i, columns 16*t + lane%16) in h2; returns y in T.y
__device__ inline void forward_tile(const MlpShared& S, WaveTile& T, f32x4 h2[4]) {
    const uint32_t lane = lane_id();
    const uint32_t r = lane % 16, q = lane / 16;
    // layer 1: [16 x 16] x [16 x 64]
    bf16x4 ax;
    for (int i = 0; i < 4; i++) ax[i] = (short)T.x[r][4 * q + i];
    f32x4 h1[4];
    for (int t = 0; t < 4; t++) {
        bf16x4 b;
        for (int i = 0; i < 4; i++) b[i] = (short)S.w1[4 * q + i][16 * t + r];
        h1[t] = mfma16(ax, b, f32x4{0, 0, 0, 0});
    }
    for (int t = 0; t < 4; t++)
        for (int i = 0; i < 4; i++) {
            const float z = h1[t][i] + S.b1[16 * t + r];
            T.h1[4 * q + i][16 * t + r] = f2bf(z > 0 ? z : 0);
        }
    wave_sync();
    // layer 2: [16 x 64] x [64 x 64]
    for (int t = 0; t < 4; t++) h2[t] = f32x4{0, 0, 0, 0};
    for (int s = 0; s < 4; s++) {
        bf16x4 a;
        for (int i = 0; i < 4; i++) a[i] = (short)T.h1[r][16 * s + 4 * q + i];
        for (int t = 0; t < 4; t++) {
            bf16x4 b;
            for (int i = 0; i < 4; i++) b[i] = (short)S.w2[16 * s + 4 * q + i][16 * t + r];
            h2[t] = mfma16(a, b, h2[t]);
        }
    }
    // bias, ReLU, layer 3 partial dots
    float part[4] = {0, 0, 0, 0};
    for (int t = 0; t < 4; t++)
        for (int i = 0; i < 4; i++) {
            const float z = h2[t][i] + S.b2[16 * t + r];
            h2[t][i] = z > 0 ? z : 0;
            part[i] += h2[t][i] * S.w3[16 * t + r];
        }
    for (int i = 0; i < 4; i++) {
        float v = part[i];
        for (int m = 8; m >= 1; m >>= 1) v += __shfl_xor(v, m, 16);
        if (r == 0) T.y[4 * q + i] = v + S.b3;
    }
    wave_sync();
}

struct MlpArgs {
    const float* __restrict__ params;   // fp32 master
    const float* __restrict__ x;        // n x 8
    const float* __restrict__ yt;       // n targets (train)
    const uint32_t* __restrict__ idx;   // batch -> sample (nullable: identity)
    uint64_t n;                         // samples in this launch
    float inv_n;                        // 1 / batch size (MSE mean)
    float* grad;                        // G2_MLP_NPARAM accumulators (train)
    float* loss;                        // sum of squared errors (train)
    float* yout;                        // predictions (infer)
    float* partial;                     // train: per wave, its gradient and squared-error sums
};

// per-wave partial sums: the G2_MLP_NPARAM gradient entries, then the sum
// of squared errors, padded to a multiple of 4 floats
constexpr uint32_t kPartialStride = (kNParam + 1 + 3) / 4 * 4;

__device__ inline void load_tile(const MlpArgs& a, WaveTile& T, uint64_t base) {
    const uint32_t lane = lane_id();
    for (uint32_t k = lane; k < 16 * 16; k += kWave) {
        const uint32_t m = k / 16, f = k % 16;
        const uint64_t s = base + m;
        float v = 0;
        if (f < kIn && s < a.n) v = a.x[(a.idx ? a.idx[s] : s) * kIn + f];
        T.x[m][f] = f2bf(v);
    }
    wave_sync();
}

__global__ __launch_bounds__(256) void mlp_infer_kernel(MlpArgs a) {
    __shared__ MlpShared S;
    __shared__ WaveTile tiles[4];
    load_weights(S, a.params);
    WaveTile& T = tiles[threadIdx.x / kWave];
    const uint32_t lane = lane_id();
    const uint64_t nt = (a.n + 15) / 16;
    for (uint64_t tile = (uint64_t)blockIdx.x * 4 + threadIdx.x / kWave; tile < nt; tile += (uint64_t)gridDim.x * 4) {
        load_tile(a, T, tile * 16);
        f32x4 h2[4];
        forward_tile(S, T, h2);
        if (lane < 16 && tile * 16 + lane < a.n) a.yout[tile * 16 + lane] = T.y[lane];
        wave_sync();
    }
}

__global__ __launch_bounds__(256) void mlp_train_kernel(MlpArgs a) {
    __shared__ MlpShared S;
    __shared__ WaveTile tiles[4];
    load_weights(S, a.params);
    WaveTile& T = tiles[threadIdx.x / kWave];
    const uint32_t lane = lane_id();
    const uint32_t r = lane % 16, q = lane / 16;
    const uint64_t nt = (a.n + 15) / 16;
    // per-wave gradient accumulators in MFMA D layout
    f32x4 gw2[4][4], gw1[4];
    for (int p = 0; p < 4; p++) {
        gw1[p] = f32x4{0, 0, 0, 0};
        for (int t = 0; t < 4; t++) gw2[p][t] = f32x4{0, 0, 0, 0};
    }
    float gw3[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb1[4] = {0, 0, 0, 0}, gb3 = 0, sse = 0;
    for (uint64_t tile = (uint64_t)blockIdx.x * 4 + threadIdx.x / kWave; tile < nt; tile += (uint64_t)gridDim.x * 4) {
        load_tile(a, T, tile * 16);
        f32x4 h2[4];
        forward_tile(S, T, h2);
        // dy for my 4 rows
        float dy[4];
        for (int i = 0; i < 4; i++) {
            const uint64_t s = tile * 16 + 4 * q + i;
            float e = 0;
            if (s < a.n) e = T.y[4 * q + i] - a.yt[a.idx ? a.idx[s] : s];
            if (r == 0) sse += e * e;
            dy[i] = 2.0f * e * a.inv_n;
            if (r == 0) gb3 += dy[i];
        }
        // dh2 = dy * w3 masked by ReLU; dW3 += h2^T dy; db2
        for (int t = 0; t < 4; t++) {
            const float w3 = S.w3[16 * t + r];
            float col3 = 0, colb = 0;
            for (int i = 0; i < 4; i++) {
                const float g = h2[t][i] > 0 ? dy[i] * w3 : 0.f;
                col3 += h2[t][i] * dy[i];
                colb += g;
                T.g[4 * q + i][16 * t + r] = f2bf(g);
            }
            gw3[t] += col3;
            gb2[t] += colb;
        }
        wave_sync();
        // dW2[k][n] += sum_m h1[m][k] dh2[m][n]   (M = k, K = samples, N = n)
        for (int p = 0; p < 4; p++) {
            bf16x4 aa;
            for (int i = 0; i < 4; i++) aa[i] = (short)T.h1[4 * q + i][16 * p + r];
            for (int t = 0; t < 4; t++) {
                bf16x4 b;
                for (int i = 0; i < 4; i++) b[i] = (short)T.g[4 * q + i][16 * t + r];
                gw2[p][t] = mfma16(aa, b, gw2[p][t]);
            }
        }
        // dh1 = dh2 W2^T masked by ReLU   (M = samples, K = n, N = k)
        f32x4 d1[4];
        for (int t = 0; t < 4; t++) d1[t] = f32x4{0, 0, 0, 0};
        for (int s = 0; s < 4; s++) {
            bf16x4 aa;
            for (int i = 0; i < 4; i++) aa[i] = (short)T.g[r][16 * s + 4 * q + i];
            for (int t = 0; t < 4; t++) {
                bf16x4 b;
                for (int i = 0; i < 4; i++) b[i] = (short)S.w2[16 * t + r][16 * s + 4 * q + i];
                d1[t] = mfma16(aa, b, d1[t]);
            }
        }
        wave_sync();
        for (int t = 0; t < 4; t++) {
            float colb = 0;
            for (int i = 0; i < 4; i++) {
                const float h = bf2f(T.h1[4 * q + i][16 * t + r]);
                const float g = h > 0 ? d1[t][i] : 0.f;
                colb += g;
                T.g[4 * q + i][16 * t + r] = f2bf(g);
            }
            gb1[t] += colb;
        }
        wave_sync();
        // dW1[k][n] += sum_m x[m][k] dh1[m][n]   (M = k padded to 16, K = samples, N = n)
        {
            bf16x4 aa;
            for (int i = 0; i < 4; i++) aa[i] = (short)T.x[4 * q + i][r];
            for (int t = 0; t < 4; t++) {
                bf16x4 b;
                for (int i = 0; i < 4; i++) b[i] = (short)T.g[4 * q + i][16 * t + r];
                gw1[t] = mfma16(aa, b, gw1[t]);
            }
        }
        wave_sync();
    }
    // flush: the wave's partial sums go to its own slot with plain stores
    // (every entry is written by exactly one lane); mlp_reduce_kernel adds
    // the slots in wave order, so training is reproducible bit for bit
    float* P = a.partial + ((uint64_t)blockIdx.x * 4 + threadIdx.x / kWave) * kPartialStride;
    for (int t = 0; t < 4; t++) {
        float v3 = gw3[t], vb2 = gb2[t], vb1 = gb1[t];
        v3 += __shfl_xor(v3, 16, 64);
        v3 += __shfl_xor(v3, 32, 64);
        vb2 += __shfl_xor(vb2, 16, 64);
        vb2 += __shfl_xor(vb2, 32, 64);
        vb1 += __shfl_xor(vb1, 16, 64);
        vb1 += __shfl_xor(vb1, 32, 64);
        if (q == 0) {
            P[oW3 + 16 * t + r] = v3;
            P[oB2 + 16 * t + r] = vb2;
            P[oB1 + 16 * t + r] = vb1;
        }
    }
    for (int p = 0; p < 4; p++)
        for (int t = 0; t < 4; t++)
            for (int i = 0; i < 4; i++) P[oW2 + (16 * p + 4 * q + i) * kH + 16 * t + r] = gw2[p][t][i];
    for (int t = 0; t < 4; t++)
        for (int i = 0; i < 4; i++) {
            const int k = 4 * q + i;
            if (k < kIn) P[oW1 + k * kH + 16 * t + r] = gw1[t][i];
        }
    for (int m = 32; m >= 1; m >>= 1) {  // lanes with r == 0 hold the partial sums
        gb3 += __shfl_xor(gb3, m, 64);
        sse += __shfl_xor(sse, m, 64);
    }
    if (lane == 0) {
        P[oB3] = gb3;
        P[kNParam] = sse;
    }
}

// grad[i] = sum over the launch's waves, in wave order, of their partial
// entry i; the squared errors are added to *loss (an epoch's running sum)
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ partial, uint32_t nwaves,
                                                         float* __restrict__ grad, float* __restrict__ loss) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > (uint32_t)kNParam) return;
    float s = 0;
    for (uint32_t w = 0; w < nwaves; w++) s += partial[(uint64_t)w * kPartialStride + i];
    if (i < (uint32_t)kNParam) grad[i] = s;
    else *loss += s;
}

struct AdamArgs {
    float* params;
    const float* grad;
    float* m;
    float* v;
    float lr, b1, b2, eps, c1, c2;  // c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
};

__global__ void adam_kernel(AdamArgs a) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kNParam; i += gridDim.x * blockDim.x) {
        const float g = a.grad[i];
        const float m = a.b1 * a.m[i] + (1 - a.b1) * g;
        const float v = a.b2 * a.v[i] + (1 - a.b2) * g * g;
        a.m[i] = m;
        a.v[i] = v;
        a.params[i] -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    }
}

}  // namespace g2

struct g2_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
    float *x = nullptr, *y = nullptr, *out = nullptr;
    float* partial = nullptr;  // per-wave partial gradients of one training launch
    uint64_t partial_waves = 0;
    uint32_t* idx = nullptr;
    uint64_t cap = 0;
    uint64_t step = 0;
    uint32_t num_cus = 256;
    std::mt19937_64 rng;
};

using namespace g2;

static int mlp_ensure(g2_mlp* M, uint64_t n) {
    if (n <= M->cap) return G2_OK;
    // drop the old buffers first; after a failed allocation every pointer is
    // either live or null, so a later call or the destroy frees each once
    float** fp[] = {&M->x, &M->y, &M->out};
    for (float** p : fp) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (M->idx) (void)hipFree(M->idx);
    M->idx = nullptr;
    M->cap = 0;
    if (hipMalloc(&M->x, n * kIn * 4) || hipMalloc(&M->y, n * 4) || hipMalloc(&M->out, n * 4) ||
        hipMalloc(&M->idx, n * 4))
        return G2_EHIP;
    M->cap = n;
    return G2_OK;
}

// one training launch over n samples (x/idx as set in the arguments):
// per-wave partial gradients, then their fixed-order reduction into grad
static int mlp_train_launch(g2_mlp* M, MlpArgs a) {
    const uint64_t blocks = std::min<uint64_t>((a.n + 63) / 64, (uint64_t)M->num_cus * 4);
    const uint64_t waves = blocks * 4;
    if (waves > M->partial_waves) {
        if (M->partial) (void)hipFree(M->partial);
        M->partial = nullptr;
        M->partial_waves = 0;
        if (hipMalloc(&M->partial, waves * kPartialStride * 4) != hipSuccess) {
            M->partial = nullptr;
            return G2_EHIP;
        }
        M->partial_waves = waves;
    }
    a.partial = M->partial;
    hipLaunchKernelGGL(mlp_train_kernel, dim3(blocks), dim3(256), 0, M->stream, a);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, M->stream, M->partial,
                       (uint32_t)waves, a.grad, a.loss);
    return hipGetLastError() == hipSuccess ? G2_OK : G2_EHIP;
}

extern "C" {

int g2_features(g2_ctx* c, int64_t* out) {
    if (!c || !out) return G2_EINVAL;
    if (!c->view_ok) return set_err(c, G2_ENOVIEW, "no view uploaded");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    return features_run(c, out);
}

int g2_core_numbers(g2_ctx* c, uint32_t* core) {
    if (!c || !core) return G2_EINVAL;
    if (!c->view_ok) return set_err(c, G2_ENOVIEW, "no view uploaded");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    int rc;
    if ((rc = ensure(c, c->counters, 8, "counters"))) return rc;
    struct CallBuf {
        g2_ctx* c;
        DevBuf b;
        ~CallBuf() {
            if (b.ptr) (void)hipFree(b.ptr);
            c->used -= b.bytes;
        }
    } buf{c, {}};
    const uint64_t nv = std::max<uint64_t>(c->nv, 1);
    if ((rc = ensure(c, buf.b, nv * 4 + 2 * ((nv + 255) & ~255ull), "core numbers"))) return rc;
    auto* dcore = (uint32_t*)buf.b.ptr;
    auto* act = (uint8_t*)buf.b.ptr + nv * 4;
    if ((rc = core_numbers_run(c, dcore, act, act + ((nv + 255) & ~255ull)))) return rc;
    if (c->nv && (e = hipMemcpy(core, dcore, c->nv * 4, hipMemcpyDeviceToHost))) return hip_err(c, e, "D2H core");
    return G2_OK;
}

int g2_root_inputs(g2_ctx* c, const uint32_t* core, float* out) {
    if (!c || !out) return G2_EINVAL;
    if (!c->view_ok) return set_err(c, G2_ENOVIEW, "no view uploaded");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    int rc;
    if ((rc = ensure(c, c->counters, 8, "counters"))) return rc;
    struct CallBuf {
        g2_ctx* c;
        DevBuf b;
        ~CallBuf() {
            if (b.ptr) (void)hipFree(b.ptr);
            c->used -= b.bytes;
        }
    } buf{c, {}};
    const uint64_t ns = std::max<uint64_t>(c->ns, 1), nv = std::max<uint64_t>(c->nv, 1);
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    const uint64_t o_x = al(ns * G2_NFEAT_INT * 8), o_core = o_x + al(ns * 8 * 4), o_act = o_core + al(nv * 4);
    if ((rc = ensure(c, buf.b, o_act + 2 * al(nv), "root inputs"))) return rc;
    char* base = (char*)buf.b.ptr;
    auto* feat = (int64_t*)base;
    auto* x = (float*)(base + o_x);
    auto* dcore = (uint32_t*)(base + o_core);
    if (core) {
        if ((e = hipMemcpyAsync(dcore, core, c->nv * 4, hipMemcpyHostToDevice, c->stream))) return hip_err(c, e, "H2D core");
    } else if ((rc = core_numbers_run(c, dcore, (uint8_t*)(base + o_act), (uint8_t*)(base + o_act + al(nv))))) {
        return rc;
    }
    if ((rc = features_dev(c, feat))) return rc;
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((c->ns + 255) / 256, 8192));
    hipLaunchKernelGGL(inputs_kernel, dim3(grid), dim3(256), 0, c->stream, feat, (const uint32_t*)c->src.ptr, dcore,
                       c->ns, c->nv, x);
    if ((e = hipGetLastError())) return hip_err(c, e, "root inputs launch");
    if (c->ns && (e = hipMemcpyAsync(out, x, c->ns * 8 * 4, hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H inputs");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "root inputs sync");
    return G2_OK;
}

int g2_mlp_create(g2_mlp** out, int device, uint64_t seed) {
    if (!out) return G2_EINVAL;
    auto* M = new g2_mlp;
    M->device = device;
    M->rng.seed(seed);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess) {
        delete M;
        return G2_EHIP;
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) == hipSuccess) M->num_cus = p.multiProcessorCount;
    if (hipMalloc(&M->params, kNParam * 4) || hipMalloc(&M->grad, kNParam * 4) || hipMalloc(&M->m, kNParam * 4) ||
        hipMalloc(&M->v, kNParam * 4) || hipMalloc(&M->loss, 4)) {
        g2_mlp_destroy(M);
        return G2_EHIP;
    }
    // He-uniform initialisation from the seed, zero biases
    std::vector<float> h(kNParam, 0.f);
    auto fill = [&](int off, int fan_in, int count) {
        const float b = std::sqrt(6.0f / fan_in);
        std::uniform_real_distribution<float> U(-b, b);
        for (int i = 0; i < count; i++) h[off + i] = U(M->rng);
    };
    fill(oW1, kIn, kIn * kH);
    fill(oW2, kH, kH * kH);
    fill(oW3, kH, kH);
    if (hipMemcpy(M->params, h.data(), kNParam * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(M->m, 0, kNParam * 4) != hipSuccess || hipMemset(M->v, 0, kNParam * 4) != hipSuccess) {
        g2_mlp_destroy(M);
        return G2_EHIP;
    }
    *out = M;
    return G2_OK;
}

void g2_mlp_destroy(g2_mlp* M) {
    if (!M) return;
    (void)hipSetDevice(M->device);
    for (void* p : {(void*)M->params, (void*)M->grad, (void*)M->m, (void*)M->v, (void*)M->loss, (void*)M->x,
                    (void*)M->y, (void*)M->out, (void*)M->idx, (void*)M->partial})
        if (p) (void)hipFree(p);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
}

int g2_mlp_get_params(g2_mlp* M, float* out) {
    if (!M || !out) return G2_EINVAL;
    return hipMemcpy(out, M->params, kNParam * 4, hipMemcpyDeviceToHost) == hipSuccess ? G2_OK : G2_EHIP;
}

int g2_mlp_set_params(g2_mlp* M, const float* in) {
    if (!M || !in) return G2_EINVAL;
    return hipMemcpy(M->params, in, kNParam * 4, hipMemcpyHostToDevice) == hipSuccess ? G2_OK : G2_EHIP;
}

int g2_mlp_infer(g2_mlp* M, const float* x, uint64_t n, float* y) {
    if (!M || (n && (!x || !y))) return G2_EINVAL;
    if (n == 0) return G2_OK;
    (void)hipSetDevice(M->device);
    if (mlp_ensure(M, n)) return G2_EHIP;
    if (hipMemcpyAsync(M->x, x, n * kIn * 4, hipMemcpyHostToDevice, M->stream)) return G2_EHIP;
    MlpArgs a{M->params, M->x, nullptr, nullptr, n, 1.f, nullptr, nullptr, M->out, nullptr};
    const uint64_t blocks = std::min<uint64_t>((n + 63) / 64, (uint64_t)M->num_cus * 4);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(256), 0, M->stream, a);
    if (hipGetLastError() != hipSuccess) return G2_EHIP;
    if (hipMemcpyAsync(y, M->out, n * 4, hipMemcpyDeviceToHost, M->stream)) return G2_EHIP;
    return hipStreamSynchronize(M->stream) == hipSuccess ? G2_OK : G2_EHIP;
}

// gradient of mean squared error over the n samples at the current parameters
int g2_mlp_grad(g2_mlp* M, const float* x, const float* y, uint64_t n, float* grad_out, float* loss_out) {
    if (!M || !n || !x || !y || !grad_out) return G2_EINVAL;
    (void)hipSetDevice(M->device);
    if (mlp_ensure(M, n)) return G2_EHIP;
    if (hipMemcpyAsync(M->x, x, n * kIn * 4, hipMemcpyHostToDevice, M->stream) ||
        hipMemcpyAsync(M->y, y, n * 4, hipMemcpyHostToDevice, M->stream) || hipMemsetAsync(M->loss, 0, 4, M->stream))
        return G2_EHIP;
    MlpArgs a{M->params, M->x, M->y, nullptr, n, 1.f / n, M->grad, M->loss, nullptr, nullptr};
    if (mlp_train_launch(M, a)) return G2_EHIP;
    float sse = 0;
    if (hipMemcpyAsync(grad_out, M->grad, kNParam * 4, hipMemcpyDeviceToHost, M->stream) ||
        hipMemcpyAsync(&sse, M->loss, 4, hipMemcpyDeviceToHost, M->stream) || hipStreamSynchronize(M->stream))
        return G2_EHIP;
    if (loss_out) *loss_out = sse / n;
    return G2_OK;
}

// Adam over shuffled mini-batches; *loss_out = mean squared error of the last epoch
int g2_mlp_train(g2_mlp* M, const float* x, const float* y, uint64_t n, uint32_t epochs, uint32_t batch,
                 float lr, float* loss_out) {
    if (!M || !n || !x || !y || !batch) return G2_EINVAL;
    (void)hipSetDevice(M->device);
    if (mlp_ensure(M, n)) return G2_EHIP;
    if (hipMemcpyAsync(M->x, x, n * kIn * 4, hipMemcpyHostToDevice, M->stream) ||
        hipMemcpyAsync(M->y, y, n * 4, hipMemcpyHostToDevice, M->stream))
        return G2_EHIP;
    std::vector<uint32_t> perm(n);
    for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
    float last = 0;
    for (uint32_t ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), M->rng);
        if (hipMemcpyAsync(M->idx, perm.data(), n * 4, hipMemcpyHostToDevice, M->stream) ||
            hipMemsetAsync(M->loss, 0, 4, M->stream))
            return G2_EHIP;
        for (uint64_t b0 = 0; b0 < n; b0 += batch) {
            const uint64_t bn = std::min<uint64_t>(batch, n - b0);
            MlpArgs a{M->params, M->x, M->y, M->idx + b0, bn, 1.f / bn, M->grad, M->loss, nullptr, nullptr};
            if (mlp_train_launch(M, a)) return G2_EHIP;
            M->step++;
            const float b1 = 0.9f, b2 = 0.999f;
            AdamArgs ad{M->params, M->grad, M->m, M->v, lr, b1, b2, 1e-8f,
                        1.f / (1.f - std::pow(b1, (float)M->step)), 1.f / (1.f - std::pow(b2, (float)M->step))};
            hipLaunchKernelGGL(adam_kernel, dim3(8), dim3(256), 0, M->stream, ad);
            if (hipGetLastError() != hipSuccess) return G2_EHIP;
        }
        float sse = 0;
        if (hipMemcpyAsync(&sse, M->loss, 4, hipMemcpyDeviceToHost, M->stream) || hipStreamSynchronize(M->stream))
            return G2_EHIP;
        last = sse / n;
    }
    if (loss_out) *loss_out = last;
    return G2_OK;
}

// The learned schedule of the uploaded view without leaving the device:
// features, the 8 inputs, their standardisation, the cost model, then the
// order (stable radix sort of the cost keys) and the parts.  The rule is
// g2h_learned_schedule's (share = max(sum of costs / waves, 1)).
int g2_learned_schedule(g2_ctx* c, g2_mlp* M, const float* mu, const float* sd, const uint32_t* core, uint64_t waves,
                        uint32_t max_split, uint32_t* order, uint32_t* split, float* pred) {
    if (!c || !M || !mu || !sd) return G2_EINVAL;
    if (!c->view_ok) return set_err(c, G2_ENOVIEW, "no view uploaded");
    if (M->device != c->device) return set_err(c, G2_EINVAL, "the cost model lives on another device");
    const uint64_t ns = c->ns, nv = std::max<uint64_t>(c->nv, 1);
    if (ns && (!order || !split)) return set_err(c, G2_EINVAL, "null order or split");
    if (ns >= 0xffffffffull) return set_err(c, G2_EINVAL, "at most 2^32-1 roots");
    if (!ns) return G2_OK;
    if (!max_split) max_split = 4096;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_err(c, e, "hipSetDevice");
    if (mlp_ensure(M, ns)) return set_err(c, G2_EHIP, "cost model buffers");
    int rc;
    if ((rc = ensure(c, c->counters, 8, "counters"))) return rc;
    struct CallBuf {
        g2_ctx* c;
        DevBuf b;
        ~CallBuf() {
            if (b.ptr) (void)hipFree(b.ptr);
            c->used -= b.bytes;
        }
    } buf{c, {}};
    constexpr uint32_t kKeyBlocks = 1024;
    const uint64_t kb = std::min<uint64_t>(kKeyBlocks, (ns + 255) / 256);
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    size_t t_sort = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, t_sort, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                (uint32_t*)nullptr, (uint32_t*)nullptr, (int)ns, 0, 32, c->stream)))
        return hip_err(c, e, "sort size");
    const uint64_t o_core = al(ns * G2_NFEAT_INT * 8), o_act = o_core + al(nv * 4), o_musd = o_act + 2 * al(nv),
                   o_keys = o_musd + 256, o_keys2 = o_keys + al(ns * 4), o_idx = o_keys2 + al(ns * 4),
                   o_order = o_idx + al(ns * 4), o_split = o_order + al(ns * 4), o_part = o_split + al(ns * 4),
                   o_temp = o_part + al(kb * 8), total = o_temp + al(t_sort);
    if ((rc = ensure(c, buf.b, total, "learned schedule"))) return rc;
    char* base = (char*)buf.b.ptr;
    auto* feat = (int64_t*)base;
    auto* dcore = (uint32_t*)(base + o_core);
    auto* musd = (float*)(base + o_musd);
    auto* keys = (uint32_t*)(base + o_keys);
    auto* keys2 = (uint32_t*)(base + o_keys2);
    auto* idx = (uint32_t*)(base + o_idx);
    auto* dorder = (uint32_t*)(base + o_order);
    auto* dsplit = (uint32_t*)(base + o_split);
    auto* part = (double*)(base + o_part);
    if (core) {
        if ((e = hipMemcpyAsync(dcore, core, c->nv * 4, hipMemcpyHostToDevice, c->stream))) return hip_err(c, e, "H2D core");
    } else if ((rc = core_numbers_run(c, dcore, (uint8_t*)(base + o_act), (uint8_t*)(base + o_act + al(nv))))) {
        return rc;
    }
    float h_musd[16];
    for (int k = 0; k < 8; k++) {
        h_musd[k] = mu[k];
        h_musd[8 + k] = sd[k];
    }
    if ((e = hipMemcpyAsync(musd, h_musd, sizeof(h_musd), hipMemcpyHostToDevice, c->stream))) return hip_err(c, e, "H2D norm");
    if ((rc = features_dev(c, feat))) return rc;  // (synchronises: h_musd and core are on the device)
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((ns + 255) / 256, 8192));
    hipLaunchKernelGGL(inputs_kernel, dim3(grid), dim3(256), 0, c->stream, (const int64_t*)feat,
                       (const uint32_t*)c->src.ptr, (const uint32_t*)dcore, ns, c->nv, M->x);
    hipLaunchKernelGGL(norm_kernel, dim3(std::min<uint64_t>((ns * 8 + 255) / 256, 8192)), dim3(256), 0, c->stream,
                       M->x, ns, (const float*)musd);
    MlpArgs a{M->params, M->x, nullptr, nullptr, ns, 1.f, nullptr, nullptr, M->out, nullptr};
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(std::min<uint64_t>((ns + 63) / 64, (uint64_t)M->num_cus * 4)), dim3(256), 0,
                       c->stream, a);
    hipLaunchKernelGGL(sched_keys_kernel, dim3(kb), dim3(256), 0, c->stream, (const float*)M->out, ns, keys, idx, part);
    if ((e = hipGetLastError())) return hip_err(c, e, "schedule launch");
    std::vector<double> h_part(kb);
    if ((e = hipMemcpyAsync(h_part.data(), part, kb * 8, hipMemcpyDeviceToHost, c->stream))) return hip_err(c, e, "D2H costs");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "schedule sync");
    double sum = 0;
    for (uint64_t b = 0; b < kb; b++) sum += h_part[b];
    const double share = std::max(sum / (double)std::max<uint64_t>(waves, 1), 1.0);
    hipLaunchKernelGGL(sched_split_kernel, dim3(grid), dim3(256), 0, c->stream, (const float*)M->out,
                       (const int64_t*)feat, ns, share, max_split, dsplit);
    size_t tb = t_sort;
    if ((e = hipcub::DeviceRadixSort::SortPairs(base + o_temp, tb, keys, keys2, idx, dorder, (int)ns, 0, 32, c->stream)))
        return hip_err(c, e, "order sort");
    if ((e = hipGetLastError())) return hip_err(c, e, "split launch");
    if ((e = hipMemcpyAsync(order, dorder, ns * 4, hipMemcpyDeviceToHost, c->stream)) ||
        (e = hipMemcpyAsync(split, dsplit, ns * 4, hipMemcpyDeviceToHost, c->stream)) ||
        (pred && (e = hipMemcpyAsync(pred, M->out, ns * 4, hipMemcpyDeviceToHost, c->stream))))
        return hip_err(c, e, "D2H schedule");
    if ((e = hipStreamSynchronize(c->stream))) return hip_err(c, e, "schedule sync");
    return G2_OK;
}

}  // extern "C"
