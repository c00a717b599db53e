// seededRandom.h — counter-based SplitMix64 streams (Steele, Lea & Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014), the
// generator every synthetic workload of this project draws from, so that
// scenes built in C++ and in Python are bit-identical.
#pragma once

#include <cstdint>
#include <vector>

constexpr uint64_t kRootSeed = 0x20261203ULL;

// n outputs of stream `stream` under `seed`: element i (1-based) is the
// SplitMix64 finaliser of base + i * golden gamma
inline std::vector<uint64_t> splitmix64(uint64_t seed, uint64_t stream, size_t n) {
    const uint64_t base = seed ^ (stream * 0xD1B54A32D192ED03ULL);
    std::vector<uint64_t> out(n);
    for (size_t i = 0; i < n; i++) {
        uint64_t z = base + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        out[i] = z ^ (z >> 31);
    }
    return out;
}

// uniform doubles in [lo, hi): the top 53 bits scaled by 2^-53
inline std::vector<double> uniform(uint64_t seed, uint64_t stream, size_t n, double lo, double hi) {
    std::vector<uint64_t> z = splitmix64(seed, stream, n);
    std::vector<double> u(n);
    for (size_t i = 0; i < n; i++) u[i] = lo + (hi - lo) * ((double)(z[i] >> 11) * 0x1.0p-53);
    return u;
}
