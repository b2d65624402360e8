// rng.h -- counter-based random streams shared by the host-side generators.
// Transaction t of a batch draws from its own SplitMix64 stream started at
// mix64(seed + (t+1) * golden), so generated batches do not depend on how
// many threads produce them.
#pragma once

#include <cstdint>

namespace ltpmg_host {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) : s(mix64(seed + kGolden * (stream + 1))) {}
    uint64_t next() {
        s += kGolden;
        return mix64(s);
    }
    // uniform in [0, n) by multiply-shift
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    // uniform in [a, b]
    uint64_t range(uint64_t a, uint64_t b) { return a + below(b - a + 1); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // TPC-C NURand(A, x, y) with a fixed C of 0
    uint64_t nurand(uint64_t A, uint64_t x, uint64_t y) {
        return ((range(0, A) | range(x, y)) % (y - x + 1)) + x;
    }
};

}  // namespace ltpmg_host
