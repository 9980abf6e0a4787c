/* Device helpers shared by the walk kernels (wave-level primitives and the
 * tally flush into 128-bit global sums). */
#ifndef TMC_DEVICE_H
#define TMC_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmc {

/* Wave-uniform copy of a 64-bit value held by the first active lane. */
__device__ __forceinline__ uint64_t rfl64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t rfl32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

/* Adds v to the 128-bit little-endian sum at p (p[0] low word, p[1] high):
 * the low word's atomic returns the old value, and a wrap-around carries
 * into the high word.  The 128-bit total is exact in any order. */
__device__ __forceinline__ void add128(unsigned long long* p, unsigned long long v)
{
    const unsigned long long old = atomicAdd(p, v);
    if (old + v < old)
        atomicAdd(p + 1, 1ull);
}

/* add128 with system-scope atomics, performed at one point for every XCD,
 * so the wave that later exchanges the sum with zero reads every wave's
 * additions. */
__device__ __forceinline__ void add128_sys(unsigned long long* p, unsigned long long v)
{
    const unsigned long long old = __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (old + v < old)
        __hip_atomic_fetch_add(p + 1, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

/* Waits until every vector memory operation this wave issued has completed
 * (s_waitcnt vmcnt(0): GFX9 encoding, expcnt and lgkmcnt left at their
 * maximum). */
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

/* Photons a wave takes from its pool [next, end) of one claimed chunk for
 * its n dead lanes: min(end - next, n).  A chunk holds fewer than 2^32
 * photons, so the count is formed in 32 bits (one scalar min; the scalar
 * unit has no 64-bit ordered compare, so a 64-bit min of wave-uniform
 * values would run on the vector unit). */
__device__ __forceinline__ uint32_t pool_take(uint64_t next, uint64_t end, uint32_t n)
{
    const uint32_t avail = (uint32_t)(end - next);
    return avail < n ? avail : n;
}

/* Rank of this lane among the lanes set in `mask` (lanes below it). */
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

} // namespace tmc

#endif
