/* Philox4x32-10 counter-based generator, for host C/C++ and HIP device code.
 *
 * Replaces the reference's glibc rand() stream (tiny_mc.c:43 srand,
 * photon.c:21,37,38,46 rand) with a generator that is a pure function of
 * (counter, key): every photon owns the stream keyed by the run seed with
 * counter (interaction k, block b, photon_lo, photon_hi), so lanes hold no
 * generator state and results do not depend on which lane walks which photon.
 *
 * Algorithm: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11 — 4 words of 32 bits, 10 rounds, multipliers 0xD2511F53 and
 * 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85.
 *
 * On gfx950 each round is two v_mad_u64_u32 (hi and lo of a 32x32 product in
 * one instruction) plus two three-input XORs; the key schedule depends only on the seed, so
 * within a wave it lives in scalar registers.
 */
#ifndef TMC_PHILOX_H
#define TMC_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define TMC_HD __host__ __device__ __forceinline__
#else
#define TMC_HD static inline
#endif

#define TMC_PHILOX_M0 0xD2511F53u
#define TMC_PHILOX_M1 0xCD9E8D57u
#define TMC_PHILOX_W0 0x9E3779B9u
#define TMC_PHILOX_W1 0xBB67AE85u

/* a ^ b ^ c.  In device code one v_bitop3_b32 (truth table 0x96, the
 * three-input exclusive or), which halves the XORs of a round: the compiler
 * emits two v_xor_b32 for the plain expression. */
TMC_HD uint32_t tmc_xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

typedef struct tmc_u32x4 {
    uint32_t x0, x1, x2, x3;
} tmc_u32x4;

TMC_HD tmc_u32x4 tmc_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)c0 * TMC_PHILOX_M0;
        const uint64_t p1 = (uint64_t)c2 * TMC_PHILOX_M1;
        const uint32_t n0 = tmc_xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = tmc_xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += TMC_PHILOX_W0;
        k1 += TMC_PHILOX_W1;
    }
    tmc_u32x4 r = { c0, c1, c2, c3 };
    return r;
}

/* The part of rounds 1-2 of block (k, 0) that depends only on the photon
 * (counter words 2-3) and the key: for counter (k, 0, c2, c3) round 1 gives
 *   c0' = hi(c2 * M1) ^ 0 ^ key0,  c1' = lo(c2 * M1),
 *   c2' = hi(k * M0) ^ c3 ^ key1,  c3' = lo(k * M0),
 * so a walk computes c0', c1' and c3 ^ key1 once per photon and only the
 * product k * M0 per absorption.  Round 2 multiplies c0' by M0, again a
 * photon constant: its high word (with round 2's key1 folded in, c2'' =
 * hi(c0' M0) ^ c3' ^ (key1 + W1)) and its low word (c3'') are kept too, for
 * block 0 (tmc_philox4x32_10_block0_r2); block b > 0 changes c0'. */
typedef struct tmc_philox_photon {
    uint32_t c0r1, c1r1, c3k1;
    uint32_t c2r2k, c3r2;
} tmc_philox_photon;

TMC_HD tmc_philox_photon tmc_philox_photon_init(uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
    const uint64_t p1 = (uint64_t)c2 * TMC_PHILOX_M1;
    const uint32_t c0r1 = (uint32_t)(p1 >> 32) ^ k0;
    const uint64_t p0 = (uint64_t)c0r1 * TMC_PHILOX_M0;
    tmc_philox_photon pp = { c0r1, (uint32_t)p1, c3 ^ k1, (uint32_t)(p0 >> 32) ^ (k1 + TMC_PHILOX_W1),
                             (uint32_t)p0 };
    return pp;
}

/* tmc_philox4x32_10(k, 0, c2, c3, k0, k1) from the photon's precomputed part. */
TMC_HD tmc_u32x4 tmc_philox4x32_10_block0(uint32_t k, tmc_philox_photon pp, uint32_t k0, uint32_t k1)
{
    const uint64_t q0 = (uint64_t)k * TMC_PHILOX_M0;
    uint32_t c0 = pp.c0r1, c1 = pp.c1r1, c2 = (uint32_t)(q0 >> 32) ^ pp.c3k1, c3 = (uint32_t)q0;
    k0 += TMC_PHILOX_W0;
    k1 += TMC_PHILOX_W1;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int round = 1; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)c0 * TMC_PHILOX_M0;
        const uint64_t p1 = (uint64_t)c2 * TMC_PHILOX_M1;
        const uint32_t n0 = tmc_xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = tmc_xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += TMC_PHILOX_W0;
        k1 += TMC_PHILOX_W1;
    }
    tmc_u32x4 r = { c0, c1, c2, c3 };
    return r;
}

/* Round keys of block (k, 0) for rounds 1-9: key0 + r W0 and key1 + r W1.
 * They depend only on the seed, so a walk computes them once per parameter
 * set and keeps them in scalar registers. */
typedef struct tmc_philox_keys {
    uint32_t k0[10], k1[10];
} tmc_philox_keys;

TMC_HD tmc_philox_keys tmc_philox_key_schedule(uint32_t k0, uint32_t k1)
{
    tmc_philox_keys ks;
    for (int r = 0; r < 10; ++r) {
        ks.k0[r] = k0 + (uint32_t)r * TMC_PHILOX_W0;
        ks.k1[r] = k1 + (uint32_t)r * TMC_PHILOX_W1;
    }
    return ks;
}

/* tmc_philox4x32_10_block0 with the round keys precomputed. */
TMC_HD tmc_u32x4 tmc_philox4x32_10_block0_ks(uint32_t k, tmc_philox_photon pp, tmc_philox_keys ks)
{
    const uint64_t q0 = (uint64_t)k * TMC_PHILOX_M0;
    uint32_t c0 = pp.c0r1, c1 = pp.c1r1, c2 = (uint32_t)(q0 >> 32) ^ pp.c3k1, c3 = (uint32_t)q0;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int round = 1; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)c0 * TMC_PHILOX_M0;
        const uint64_t p1 = (uint64_t)c2 * TMC_PHILOX_M1;
        const uint32_t n0 = tmc_xor3((uint32_t)(p1 >> 32), c1, ks.k0[round]);
        const uint32_t n2 = tmc_xor3((uint32_t)(p0 >> 32), c3, ks.k1[round]);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    tmc_u32x4 r = { c0, c1, c2, c3 };
    return r;
}

/* tmc_philox4x32_10_block0_ks with round 2's product of the photon constant
 * c0' taken from pp: one multiplication fewer per block (two more words per
 * photon). */
TMC_HD tmc_u32x4 tmc_philox4x32_10_block0_r2(uint32_t k, tmc_philox_photon pp, tmc_philox_keys ks)
{
    const uint64_t q0 = (uint64_t)k * TMC_PHILOX_M0;
    /* after round 1: (c0r1, c1r1, hi(q0) ^ c3k1, lo(q0)); round 2: */
    const uint64_t p1 = (uint64_t)((uint32_t)(q0 >> 32) ^ pp.c3k1) * TMC_PHILOX_M1;
    uint32_t c0 = tmc_xor3((uint32_t)(p1 >> 32), pp.c1r1, ks.k0[1]), c1 = (uint32_t)p1,
             c2 = pp.c2r2k ^ (uint32_t)q0, c3 = pp.c3r2;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int round = 2; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)c0 * TMC_PHILOX_M0;
        const uint64_t p1 = (uint64_t)c2 * TMC_PHILOX_M1;
        const uint32_t n0 = tmc_xor3((uint32_t)(p1 >> 32), c1, ks.k0[round]);
        const uint32_t n2 = tmc_xor3((uint32_t)(p0 >> 32), c3, ks.k1[round]);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    tmc_u32x4 r = { c0, c1, c2, c3 };
    return r;
}

/* Block (k, b) from the photon's part of round 1: counter word 1 enters
 * round 1 only through c0' = hi(c2 M1) ^ b ^ key0, so block b costs what
 * block 0 does plus one exclusive or.  Round keys from `ks` (STORED) or
 * formed from key0/key1. */
TMC_HD tmc_u32x4 tmc_philox4x32_10_blockb(uint32_t k, uint32_t b, tmc_philox_photon pp, const tmc_philox_keys* ks,
                                          uint32_t k0, uint32_t k1)
{
    const uint64_t q0 = (uint64_t)k * TMC_PHILOX_M0;
    uint32_t c0 = pp.c0r1 ^ b, c1 = pp.c1r1, c2 = (uint32_t)(q0 >> 32) ^ pp.c3k1, c3 = (uint32_t)q0;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int round = 1; round < 10; ++round) {
        const uint32_t rk0 = ks ? ks->k0[round] : k0 + (uint32_t)round * TMC_PHILOX_W0;
        const uint32_t rk1 = ks ? ks->k1[round] : k1 + (uint32_t)round * TMC_PHILOX_W1;
        const uint64_t p0 = (uint64_t)c0 * TMC_PHILOX_M0;
        const uint64_t p1 = (uint64_t)c2 * TMC_PHILOX_M1;
        const uint32_t n0 = tmc_xor3((uint32_t)(p1 >> 32), c1, rk0);
        const uint32_t n2 = tmc_xor3((uint32_t)(p0 >> 32), c3, rk1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    tmc_u32x4 r = { c0, c1, c2, c3 };
    return r;
}

/* 32 random bits -> uniform in (0, 1]: ((r >> 8) + 1) * 2^-24, exact in
 * binary32.  The reference maps rand() to [0, 1] (photon.c:21); excluding 0
 * keeps -log(xi) finite, which the reference does not guarantee. */
TMC_HD float tmc_u01(uint32_t r)
{
    return (float)((r >> 8) + 1u) * 0x1p-24f;
}

/* The same uniform times 2^24: the integer (r >> 8) + 1 as a float. */
TMC_HD float tmc_u24(uint32_t r)
{
    return (float)((r >> 8) + 1u);
}

/* tmc_u01(r) from one conversion and one FMA: (r >> 8) 2^-24 + 2^-24 is
 * exact in binary32 (at most 24 significant bits), so the FMA's single
 * rounding returns the same float. */
TMC_HD float tmc_u01_fma(uint32_t r)
{
    return __builtin_fmaf((float)(r >> 8), 0x1p-24f, 0x1p-24f);
}

/* 2u - 1 for the same uniform: (r >> 8) 2^-23 + (2^-23 - 1), an exact
 * multiple of 2^-23 in (-1, 1], equal to fma(tmc_u24(r), 2^-23, -1). */
TMC_HD float tmc_u11(uint32_t r)
{
    return __builtin_fmaf((float)(r >> 8), 0x1p-23f, 0x1p-23f - 1.0f);
}

#endif
