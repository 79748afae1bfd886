// csrc/g2_device.hpp -- wave64 primitives for gfx950 (CDNA4).
//
// Every search kernel in this library runs one task per 64-lane wavefront.
// These helpers replace the reference's 32-lane idioms (system/util.h:7-55,
// common/gpu_env.h:30-32): 64-bit ballots, mbcnt-based exclusive prefix,
// readfirstlane broadcasts and xor-shuffle reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace g2 {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// 64-bit ballot of a predicate over the active lanes
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set predicate bits in lanes below this one (exclusive prefix)
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// broadcast of lane 0's value (values that are already wave-uniform)
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uniform(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// value of lane `src` (src wave-uniform)
__device__ __forceinline__ uint32_t read_lane(uint32_t x, uint32_t src) { return __builtin_amdgcn_readlane(x, src); }
__device__ __forceinline__ uint64_t read_lane(uint64_t x, uint32_t src) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, src);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_xor(uint64_t x, int m) {
    uint32_t lo = __shfl_xor((uint32_t)x, m, kWave);
    uint32_t hi = __shfl_xor((uint32_t)(x >> 32), m, kWave);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_max(uint64_t x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t y = shfl_xor(x, m);
        x = y > x ? y : x;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x += shfl_xor(x, m);
    return x;
}

__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x |= shfl_xor(x, m);
    return x;
}

// Order the memory operations of all lanes of this wave: what any lane wrote
// before is visible to every lane after (global and LDS, workgroup scope).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Next task of a persistent loop: lane 0 takes it from the counter and the
// whole wave gets it.  The wave barrier (a convergent operation) pins the
// fetch at a point where the wave is converged.  Without it the compiler may
// thread a "lane 0 only" branch at the end of one iteration (a result store)
// into the next iteration's "lane 0 only" fetch; the structurised loop then
// lets lanes 1..63 go round without passing the fetch, reading a stale index
// forever.
__device__ __forceinline__ uint64_t fetch_task(unsigned long long* ctr) {
    __builtin_amdgcn_wave_barrier();
    unsigned long long t = 0;
    if (lane_id() == 0) t = atomicAdd(ctr, 1ull);
    return uniform((uint64_t)t);
}

// First of k consecutive tasks (same wave barrier as fetch_task).  One
// device-scope atomic per fetch: a counter takes roughly one atomic per 10 ns,
// so millions of cheap tasks fetched one at a time queue behind it.
__device__ __forceinline__ uint64_t fetch_tasks(unsigned long long* ctr, uint32_t k) {
    __builtin_amdgcn_wave_barrier();
    unsigned long long t = 0;
    if (lane_id() == 0) t = atomicAdd(ctr, (unsigned long long)k);
    return uniform((uint64_t)t);
}

// first index in [lo,hi) whose value is >= val (hi if none)
__device__ __forceinline__ uint64_t lower_bound(const uint32_t* __restrict__ cols, uint64_t lo, uint64_t hi, uint32_t val) {
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (cols[mid] < val) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// integer min / max (explicit template arguments to HIP's min<T> select its
// floating-point template, which converts through f64)
__host__ __device__ __forceinline__ uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ uint32_t umax32(uint32_t a, uint32_t b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

// lower bounds of x and of y in one sorted range, searched in lockstep so
// that the two loads of every step are in flight together
__device__ __forceinline__ void lower_bound2(const uint32_t* __restrict__ cols, uint64_t lo, uint64_t hi, uint32_t x,
                                             uint32_t y, uint64_t* px, uint64_t* py) {
    const uint64_t n = hi - lo;
    uint64_t cx = 0, cy = 0, step = 1;  // counts of entries below x, below y
    while (step * 2 <= n) step *= 2;
    for (step = n ? step : 0; step; step >>= 1) {
        const bool gx = cx + step <= n && cols[lo + cx + step - 1] < x;
        const bool gy = cy + step <= n && cols[lo + cy + step - 1] < y;
        cx += gx ? step : 0;
        cy += gy ? step : 0;
    }
    *px = lo + cx;
    *py = lo + cy;
}

// membership of val in the sorted row cols[lo, hi); an empty range is
// "absent" (system/util.h:19-30 reads cols[lo] there instead)
__device__ __forceinline__ bool contains(const uint32_t* __restrict__ cols, uint64_t lo, uint64_t hi, uint32_t val) {
    uint64_t p = lower_bound(cols, lo, hi, val);
    return p < hi && cols[p] == val;
}

__device__ __forceinline__ uint64_t bitlen(uint64_t d) { return d == 0 ? 0 : 64 - __builtin_clzll(d); }

// modelled bytes of intersecting a list of a entries with a row of d entries:
// the cheaper of merge streaming and per-entry binary search
__device__ __forceinline__ uint64_t cost_isect(uint64_t a, uint64_t d) {
    uint64_t merge = a + d, bs = a * (1 + bitlen(d));
    return 4 * (merge < bs ? merge : bs);
}

}  // namespace g2
