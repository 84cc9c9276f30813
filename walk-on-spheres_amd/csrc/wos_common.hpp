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
// easy as 1, 2, 3", SC'11): the four 32-bit words of the block at counter
// {c0, c1, c2, c3} under key {k0, k1}.
struct Words4 { uint32_t w0, w1, w2, w3; };

WOS_HD Words4 philox4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
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
    return Words4{c0, c1, c2, c3};
}

WOS_HD uint32_t word_at(const Words4 &b, uint32_t k) {
    return k == 0 ? b.w0 : k == 1 ? b.w1 : k == 2 ? b.w2 : b.w3;
}

// The counter stream of one walk step (walk w, step s, key = the point's
// walk-stream seed).  Its uniforms are numbered by slot k = 0, 1, 2, ...:
// uniform k takes its top 32 bits from word k mod 4 of the block at counter
// {w, s, k / 4, 0} and, in FP64, its next 21 bits from the top of the same
// word of the block at {w, s, k / 4, 1}.  So an FP32 draw is exactly the
// top 24 bits of the FP64 draw of the same slot (common random numbers
// between the walkers), and one block serves four FP32 uniforms.
// Slots 0 .. dim-2 are the step direction; the source term draws from slot
// dim-1 on (its rejection samplers continue into later blocks).
WOS_HD uint64_t bits53(uint32_t hi, uint32_t lo) { return ((uint64_t)hi << 21) | (lo >> 11); }
// uniform in [0, 1): k 2^-53 (include/wos/prng.hpp:25-27 scaling)
WOS_HD double u53(uint64_t x) { return (double)x * (1.0 / 9007199254740992.0); }
// uniform in (0, 1): the odd multiple (2j+1) 2^-53 of the cell midpoint
WOS_HD double u53_open(uint64_t x) { return (double)(x | 1ULL) * (1.0 / 9007199254740992.0); }
// FP32 forms from the top 24 bits of the slot's high word
WOS_HD float u24(uint32_t hi) { return (float)(hi >> 8) * (1.0f / 16777216.0f); }
WOS_HD float u24_open(uint32_t hi) { return (float)((hi >> 8) | 1u) * (1.0f / 16777216.0f); }

// Sequential FP64 draws of a step stream from slot `first` on.
struct StepStreamD {
    static constexpr bool radius_first = true;  // a ball point draws its radius first
    uint32_t walk, step, k0, k1, next, held;  // held: block index + 1 in hi/lo (0: none)
    Words4 hi, lo;
    WOS_HD StepStreamD(uint32_t w, uint32_t s, uint32_t key0, uint32_t key1, uint32_t first)
        : walk(w), step(s), k0(key0), k1(key1), next(first), held(0), hi{}, lo{} {}
    WOS_HD uint64_t bits() {
        const uint32_t k = next++, blk = k >> 2;
        if (held != blk + 1) {
            hi = philox4(walk, step, blk, 0, k0, k1);
            lo = philox4(walk, step, blk, 1, k0, k1);
            held = blk + 1;
        }
        return bits53(word_at(hi, k & 3), word_at(lo, k & 3));
    }
    WOS_HD double unit() { return u53(bits()); }
    WOS_HD double unit_open() { return u53_open(bits()); }
    WOS_HD void seek(uint32_t slot) { next = slot; }  // the held block stays valid
};

// Sequential FP32 draws of a step stream from slot `first` on; `b0` may
// hand over block 0, already formed by the caller.
struct StepStreamF {
    static constexpr bool radius_first = true;
    uint32_t walk, step, k0, k1, next, held;
    Words4 hi;
    WOS_HD StepStreamF(uint32_t w, uint32_t s, uint32_t key0, uint32_t key1, uint32_t first)
        : walk(w), step(s), k0(key0), k1(key1), next(first), held(0), hi{} {}
    WOS_HD StepStreamF(uint32_t w, uint32_t s, uint32_t key0, uint32_t key1, uint32_t first,
                       const Words4 &b0)
        : walk(w), step(s), k0(key0), k1(key1), next(first), held(1), hi(b0) {}
    WOS_HD uint32_t word() {
        const uint32_t k = next++, blk = k >> 2;
        if (held != blk + 1) {
            hi = philox4(walk, step, blk, 0, k0, k1);
            held = blk + 1;
        }
        return word_at(hi, k & 3);
    }
    WOS_HD float unit() { return u24(word()); }
    WOS_HD float unit_open() { return u24_open(word()); }
};

}  // namespace wosx
