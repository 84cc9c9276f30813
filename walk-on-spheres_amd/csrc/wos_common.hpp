// wos_common.hpp -- host/device integer kernels shared by the host runtime
// and the HIP kernels: the reference's seed derivation (bit-exact) and the
// counter-based walk streams.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define WOS_HD __host__ __device__ __forceinline__
#else
#define WOS_HD inline
#endif

namespace wosx {

constexpr double kPi = 3.141592653589793238462643383279502884;

// Stream constants of the reference sweep (include/wos/runner.hpp:182, :208, :211).
constexpr uint64_t kInsideStream = 0x243F6A8885A308D3ULL;
constexpr uint64_t kWalkStream = 0x13198A2E03707344ULL;
constexpr uint64_t kSourceStream = 0xA4093822299F31D0ULL;  // 2D only (runner.hpp:211)

// include/wos/hash.hpp:7-12
WOS_HD uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// include/wos/runner.hpp:163-168: seed of the point with global linear index
WOS_HD uint64_t point_seed(uint64_t seed, uint64_t linear_index) {
    return splitmix64(seed ^ splitmix64(linear_index));
}

WOS_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256++ with the reference's seeding (src/prng.cpp:6-15,
// include/wos/prng.hpp:12-27).  Used for the 3D inside-test rays, whose
// draws must equal the reference's to reproduce its location codes.
struct Xoshiro {
    uint64_t s0, s1, s2, s3;
    WOS_HD Xoshiro() : s0(0), s1(0), s2(0), s3(0) {}
    WOS_HD explicit Xoshiro(uint64_t seed) {
        s0 = splitmix64(seed);
        s1 = splitmix64(seed + 0x9E3779B97F4A7C15ULL);
        s2 = splitmix64(seed + 0x3C6EF372FE94F82AULL);
        s3 = splitmix64(seed + 0xDAA66D2C7DDF743FULL);
        if (!(s0 | s1 | s2 | s3)) s0 = 1;
    }
    WOS_HD uint64_t next() {
        const uint64_t result = rotl64(s0 + s3, 23) + s0;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl64(s3, 45);
        return result;
    }
    WOS_HD double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // include/wos/prng.hpp:31-37: redraw zero
    WOS_HD double unit_open() {
        double v;
        do { v = unit(); } while (v == 0.0);
        return v;
    }
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Counter {walk, step, block, tag}, key = the
// point's walk-stream seed.  Two u64 per call.
struct U64x2 { uint64_t a, b; };

WOS_HD U64x2 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    return U64x2{(uint64_t)c0 | ((uint64_t)c1 << 32), (uint64_t)c2 | ((uint64_t)c3 << 32)};
}

// uniform in [0, 1) from the top 53 bits (include/wos/prng.hpp:25-27)
WOS_HD double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }
// uniform in (0, 1): midpoint of a 2^-52 cell, (2k+1) 2^-53
WOS_HD double u01_open(uint64_t x) {
    return (double)(((x >> 12) << 1) | 1ULL) * (1.0 / 9007199254740992.0);
}
// FP32 forms from the top 24 bits of the same word (common random numbers
// between the FP32 and FP64 walkers)
WOS_HD float u01f(uint64_t x) { return (float)(uint32_t)(x >> 40) * (1.0f / 16777216.0f); }
WOS_HD float u01f_open(uint64_t x) {
    return (float)(uint32_t)(((x >> 41) << 1) | 1ULL) * (1.0f / 16777216.0f);
}

// Draw cursor over the counter stream of one (walk, step, tag).
struct Cursor {
    uint32_t walk, step, block, tag, k0, k1;
    uint64_t pending;
    bool has;
    WOS_HD Cursor(uint32_t w, uint32_t s, uint32_t t, uint32_t key0, uint32_t key1)
        : walk(w), step(s), block(0), tag(t), k0(key0), k1(key1), pending(0), has(false) {}
    WOS_HD uint64_t next() {
        if (has) { has = false; return pending; }
        U64x2 r = philox(walk, step, block++, tag, k0, k1);
        pending = r.b;
        has = true;
        return r.a;
    }
    WOS_HD double unit() { return u01(next()); }
    WOS_HD double unit_open() { return u01_open(next()); }
    WOS_HD float unitf() { return u01f(next()); }
    WOS_HD float unitf_open() { return u01f_open(next()); }
};

}  // namespace wosx
