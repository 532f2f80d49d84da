// lt_rng.h — counter-based random draws shared by host and device code.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): a bijection of a 128-bit counter under a 64-bit key.
// Every random decision of a walk is a pure function of (seed, walker id,
// step, tag), so results do not depend on which thread or which scheduler
// iteration performs the step.  This replaces the reference's per-thread
// curand XORWOW states (src/Rand.h:9-37).
//
// All decisions are integer-only:
//   * a bounded draw in [0, n) is the high 64 bits of r64 * n (multiply-shift),
//     replacing the reference's ((u64)curand<<32 + curand) % n
//     (src/Rand.h:17-19); the bias is at most n / 2^64;
//   * a Bernoulli(p) test is r32 < floor(p * 2^32), replacing
//     curand_uniform_double(&state) compared with alpha or prob.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define LT_HD __host__ __device__ __forceinline__
#else
#define LT_HD static inline
#endif

namespace lt {

struct u32x4 { uint32_t x, y, z, w; };

// Stream tags: the fourth counter word separates independent draw streams.
enum : uint32_t {
    TAG_START = 0x53545254u,   // start vertex of a uniform-start walker
    TAG_PPRLEN = 0x504C454Eu,  // geometric length of a PPR walker
    TAG_PAGERANK = 0x50524B53u,// pagerank step
    TAG_PPR = 0x50505253u,     // ppr step
    TAG_GENERIC = 0x47454E53u, // genericwalk step
    TAG_RMAT = 0x524D4154u,    // R-MAT quadrant choices
    TAG_QUERY = 0x51525953u    // batched-PPR query table
};

LT_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

LT_HD u32x4 draw(uint64_t seed, uint64_t id, uint32_t ctr, uint32_t tag) {
    u32x4 c{(uint32_t)id, (uint32_t)(id >> 32), ctr, tag};
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// High 64 bits of a * b.
LT_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Uniform integer in [0, n) from two 32-bit words (hi, lo).
LT_HD uint64_t bounded(uint32_t hi, uint32_t lo, uint64_t n) {
    return mulhi64(((uint64_t)hi << 32) | lo, n);
}

// floor(p * 2^32) clamped to [0, 2^32-1]; a draw r32 < threshold(p) has
// probability threshold(p)/2^32.
static inline uint32_t prob_threshold(double p) {
    if (!(p > 0.0)) return 0u;
    if (p >= 1.0) return 0xFFFFFFFFu;
    double t = p * 4294967296.0;
    return t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t;
}

}  // namespace lt
