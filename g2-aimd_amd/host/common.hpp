// host/common.hpp -- shared internals of libg2host.so
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "g2host.h"

struct g2h_graph {
    uint64_t nv = 0;
    uint64_t ne = 0;
    std::vector<uint64_t> rp;   // nv + 1 row offsets
    std::vector<uint32_t> cols; // ne sorted column ids
};

namespace g2h {

// per-thread last error
void set_error(const std::string& msg);
// progress line on stderr for long host steps (off unless g2h_set_progress(1))
void progress(const char* fmt, ...);

// splitmix64-seeded xoshiro256** (Blackman & Vigna, "Scrambled linear
// pseudorandom number generators", 2018)
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) {
            seed += 0x9e3779b97f4a7c15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0,1) from the top 53 bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0,n) (Lemire's multiply-shift, bias < 2^-64*n)
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

// build a CSR from an undirected edge list: symmetrise, sort rows, drop self
// loops and parallel edges; keeps every vertex id in [0,nv)
void csr_from_edges(uint64_t nv, const std::vector<uint32_t>& a, const std::vector<uint32_t>& b,
                    g2h_graph& g, int threads);

}  // namespace g2h
