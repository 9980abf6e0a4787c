/* Oracle-side Philox4x32-10 (test infrastructure; never linked into the product).
 *
 * Counter-based generator of Salmon, Moraes, Dror and Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3" (SC'11), 4x32 variant, 10 rounds.  Written
 * independently of the product's device header so that the two can be checked
 * against each other and against the published known-answer vectors.
 *
 * The reference (photon.c:21,37,38,46 via tiny_mc.c:43) draws from glibc
 * rand(); the product replaces that stream with this generator keyed by
 * (seed, photon index), so this file is the oracle for the replacement RNG.
 */
#ifndef ORC_PHILOX_H
#define ORC_PHILOX_H

#include <stdint.h>

#define ORC_PHILOX_M0 0xD2511F53u
#define ORC_PHILOX_M1 0xCD9E8D57u
#define ORC_PHILOX_W0 0x9E3779B9u
#define ORC_PHILOX_W1 0xBB67AE85u

static inline void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)ORC_PHILOX_M0 * c0;
        const uint64_t p1 = (uint64_t)ORC_PHILOX_M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += ORC_PHILOX_W0;
        k1 += ORC_PHILOX_W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Block b of interaction k of photon p under a 64-bit seed:
 * counter = (k, b, p_lo, p_hi), key = (seed_lo, seed_hi). */
static inline void orc_block(uint64_t seed, uint64_t photon, uint32_t k, uint32_t b, uint32_t out[4])
{
    const uint32_t ctr[4] = { k, b, (uint32_t)photon, (uint32_t)(photon >> 32) };
    const uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    orc_philox4x32_10(ctr, key, out);
}

/* 32 random bits -> uniform in (0, 1]: ((r >> 8) + 1) * 2^-24, exact in binary32. */
static inline float orc_u01(uint32_t r)
{
    return (float)((r >> 8) + 1u) * 0x1p-24f;
}

#endif
