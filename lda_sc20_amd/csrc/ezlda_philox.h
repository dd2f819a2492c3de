// Counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), usable from host and device code.
//
// Replaces the clock-seeded cuRAND XORWOW state of the reference
// (EZLDA/utility.cu:1209-1213 initRandState, drawn at utility.cu:1353 and
// 2848-2851): a draw is a pure function of (seed; token, sweep, tag), so no
// generator state lives in memory and every sweep is reproducible.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define EZ_HD __host__ __device__ __forceinline__
#else
#define EZ_HD static inline
#endif

// Counter word 3 tags one stream of draws from another.
enum {
    EZ_TAG_SAMPLE = 0u,   // one uniform per token per training sweep
    EZ_TAG_INIT = 1u,     // initial topic of a token
    EZ_TAG_GEN = 2u,      // synthetic corpus generator
    EZ_TAG_FOLDIN = 3u,   // fold-in sampling
    EZ_TAG_FOLDIN_INIT = 4u,
};

struct ez_u4 { uint32_t x, y, z, w; };

EZ_HD uint32_t ez_mulhi32(uint32_t a, uint32_t b) {
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

EZ_HD ez_u4 ez_philox4x32_10(ez_u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = ez_mulhi32(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = ez_mulhi32(0xCD9E8D57u, c.z);
        ez_u4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// First output word of the block keyed by seed with counter
// (token lo, token hi, sweep, tag).
EZ_HD uint32_t ez_philox_x0(uint64_t seed, uint64_t token, uint32_t sweep, uint32_t tag) {
    ez_u4 c;
    c.x = (uint32_t)token;
    c.y = (uint32_t)(token >> 32);
    c.z = sweep;
    c.w = tag;
    return ez_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32)).x;
}

// u in [0,1) on the 2^-24 grid: exact in f32, never 1.0.
EZ_HD float ez_u24(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }
