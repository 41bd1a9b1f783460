// Counter-seeded pseudo-random streams for the seeded matrix generators.
//
// Every independent unit of work (a founder, a row, a block of edges) gets its
// own stream seeded from (seed, stream id) through splitmix64, so generated
// matrices do not depend on the thread count. The generator is xoshiro256**
// (Blackman and Vigna, "Scrambled linear pseudorandom number generators", 2021);
// doubles take the top 53 bits.
#pragma once

#include <cmath>
#include <cstdint>

namespace gmr {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline uint64_t mix64(uint64_t x) {
    uint64_t s = x;
    return splitmix64(s);
}

struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream) {
        uint64_t z = seed ^ mix64(stream + 0x632BE59BD9B4E019ULL);
        for (auto& w : s) w = splitmix64(z);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
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
    // Uniform in [0, 1).
    inline double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    // Uniform integer in [0, n), n >= 1 (Lemire's multiply-shift, bias < n/2^64).
    inline uint64_t below(uint64_t n) {
        return static_cast<uint64_t>((static_cast<unsigned __int128>(next()) * n) >> 64);
    }
    // Geometric on {1, 2, ...} with the given mean (success probability 1/mean).
    inline long long geometric(double mean) {
        if (mean <= 1.0) return 1;
        double u = 1.0 - uniform();  // (0, 1]
        return 1 + static_cast<long long>(std::floor(std::log(u) / std::log1p(-1.0 / mean)));
    }
};

}  // namespace gmr
