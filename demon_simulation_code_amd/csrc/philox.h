// philox.h -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable on the
// host and on gfx950.  Replaces the reference's mt19937_64 RandomNumbers
// (RandomNumbers.cpp:19-50), whose wall-clock seed makes draws unrepeatable,
// with draws keyed by (seed, particle, step, stage).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define DEMON_HD __host__ __device__ __forceinline__
#else
#define DEMON_HD inline
#endif

namespace demon {

struct U32x4 {
    uint32_t v[4];
};

DEMON_HD U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                              uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    U32x4 o;
    o.v[0] = c0;
    o.v[1] = c1;
    o.v[2] = c2;
    o.v[3] = c3;
    return o;
}

// 53 random bits of (hi:lo) as a double in [0, 1).
DEMON_HD double u01(uint32_t lo, uint32_t hi) {
    const uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)v * 0x1.0p-53;
}

// Counter domains (top byte of the fourth counter word).
enum : uint32_t {
    kDomainThermal = 1u,
    kDomainLattice = 2u,
    kDomainThermalLocalized = 3u,
    kDomainTimeVaryingThermal = 4u
};

DEMON_HD U32x4 draw(uint64_t seed, uint64_t particle, uint64_t step, uint32_t domain,
                    uint32_t stage) {
    return philox4x32_10((uint32_t)particle, (uint32_t)(particle >> 32), (uint32_t)step,
                         (domain << 24) | (stage << 16) | ((uint32_t)(step >> 32) & 0xFFFFu),
                         (uint32_t)seed, (uint32_t)(seed >> 32));
}

}  // namespace demon
