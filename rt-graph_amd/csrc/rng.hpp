// Counter-based random numbers shared by the host and device generators.
//
// Every draw is a pure function of (seed, stream, counter), so a generator gives the same
// output whether it runs on one host thread, many OpenMP threads or a GPU grid. The mixer is
// the SplitMix64 finalizer (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014), applied twice so that neighbouring counters decorrelate.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define RTG_HD __host__ __device__ __forceinline__
#else
#define RTG_HD inline
#endif

namespace rtg {

RTG_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One 64-bit draw for (seed, stream, ctr).
RTG_HD uint64_t crand(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return mix64(mix64(seed * 0xD6E8FEB86659FD93ull + stream) ^ (ctr * 0x9E3779B97F4A7C15ull));
}

// Uniform integer in [0, bound) from the top 32 bits (multiply-high, bound < 2^32).
RTG_HD uint32_t bounded32(uint64_t r, uint32_t bound) {
    return (uint32_t)(((r >> 32) * (uint64_t)bound) >> 32);
}

// Uniform 53-bit integer; compare against thresholds made with prob53().
RTG_HD uint64_t top53(uint64_t r) { return r >> 11; }

inline uint64_t prob53(double p) {
    if (p <= 0.0) return 0;
    if (p >= 1.0) return 1ull << 53;
    return (uint64_t)(p * 9007199254740992.0);  // p * 2^53
}

RTG_HD double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Keyed bijection on [0, 2^bits): rounds of odd multiply, add and xor-shift, each of which is
// invertible modulo 2^bits. Used as the Graph500-style vertex relabelling of R-MAT.
RTG_HD uint32_t permute_bits(uint32_t x, uint32_t bits, uint64_t seed) {
    if (bits == 0) return 0;
    const uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
    uint64_t v = x & mask;
    const uint32_t sh = bits / 2 + 1;
    for (int round = 0; round < 4; ++round) {
        uint64_t k = mix64(seed + 0x632BE59BD9B4E019ull * (uint64_t)(round + 1));
        v = (v * (k | 1ull)) & mask;
        v = (v + (k >> 17)) & mask;
        v ^= (v >> sh);
    }
    return (uint32_t)v;
}

// Stream identifiers (one per generator so that configs never share draws).
enum : uint64_t {
    kStreamRmat = 0x524D4154ull,        // "RMAT"
    kStreamRmatPerm = 0x5045524Dull,    // "PERM"
    kStreamBA = 0x42410000ull,          // "BA"
    kStreamDegree = 0x44454700ull,      // "DEG"
    kStreamStub = 0x53545542ull,        // "STUB"
    kStreamClose = 0x434C4F53ull,       // "CLOS"
    kStreamZipf = 0x5A495046ull,        // "ZIPF"
    kStreamElem = 0x454C454Dull,        // "ELEM"
    kStreamArray = 0x41525241ull,       // "ARRA"
    kStreamQuery = 0x51554552ull,       // "QUER"
};

}  // namespace rtg
