/* The photon random walk as a persistent gfx950 kernel.
 *
 * Reference: photon() (photon.c:6-51) called PHOTONS times by tiny_mc.c:47-49.
 * Here one launch walks the whole photon range:
 *
 *  - Persistent grid with per-lane regeneration.  Each lane walks one photon
 *    at a time; a lane whose photon dies (roulette, photon.c:45-49) takes the
 *    next photon index from a pool its wave claims from a global counter
 *    `chunk` photons at a time, so wavefronts stay full until the range runs
 *    out.  Every wave leaves the loop once the range is exhausted and its
 *    lanes are dead.
 *  - Counter-based randomness: interaction k of photon p reads Philox block
 *    (k, b, p_lo, p_hi) under key = seed.  Block 0 gives the step (word 0),
 *    the direction (words 1-2) and the roulette draw (word 3); rejection
 *    attempts past the first read blocks b = 1, 2, ... two attempts each.
 *  - Tallies: each workgroup keeps 64-bit fixed-point per-shell sums of w and
 *    w*w in LDS (ds_add_u64), flushed into 128-bit global sums by atomic
 *    swap, so no barrier is needed after the initial clear and the integer
 *    totals are independent of order and launch shape.
 *  - FP32 transport: step, position, radius, direction and weight in binary32
 *    with fused multiply-adds; log/sin/cos from tmc_math.h; IEEE sqrt and
 *    division.  The host mirror reproduces every operation bit for bit.
 *
 * No MFMA: the walk is a scalar, transcendental-heavy loop with no dense
 * contraction; it is bound by VALU issue.
 */
#include "tmc_internal.h"
#include "tmc_math.h"
#include "tmc_philox.h"

namespace tmc {

__device__ __forceinline__ uint64_t rfl64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ void add128(unsigned long long* p, unsigned long long v)
{
    const unsigned long long old = atomicAdd(p, v);
    if (old + v < old)
        atomicAdd(p + 1, 1ull);
}

/* Moves this workgroup's LDS tallies into the global 128-bit sums.  Safe
 * while other waves of the workgroup keep adding: each entry is taken by an
 * atomic swap with zero, so every addition is counted exactly once. */
__device__ __forceinline__ void flush_tallies(unsigned long long* lw, unsigned long long* lw2,
                                              unsigned long long* g, uint32_t shells, uint32_t lane)
{
    for (uint32_t s = lane; s < shells; s += 64) {
        const unsigned long long a = atomicExch(&lw[s], 0ull);
        const unsigned long long b = atomicExch(&lw2[s], 0ull);
        if (a)
            add128(g + 2 * s, a);
        if (b)
            add128(g + 2 * shells + 2 * s, b);
    }
}

template <uint32_t SAMPLER>
__global__ void __launch_bounds__(1024) walk_f32(WalkArgs a)
{
    extern __shared__ unsigned long long lds_tally[];
    unsigned long long* lw = lds_tally;
    unsigned long long* lw2 = lds_tally + a.shells;
    for (uint32_t i = threadIdx.x; i < 2 * a.shells; i += blockDim.x)
        lds_tally[i] = 0;
    __syncthreads();

    const uint32_t lane = __lane_id();
    const uint32_t key0 = (uint32_t)a.seed;
    const uint32_t key1 = (uint32_t)(a.seed >> 32);

    /* wave-uniform pool of claimed photon offsets [pool_next, pool_end) */
    uint64_t pool_next = 0, pool_end = 0;
    bool exhausted = false;
    uint64_t wave_interactions = 0, wave_truncated = 0;
    uint32_t since_flush = 0;

    /* lane state: one photon (photon.c:11-18) */
    bool alive = false;
    float x = 0.0f, y = 0.0f, z = 0.0f, u = 0.0f, v = 0.0f, w = 1.0f, weight = 1.0f;
    uint32_t k = 0, plo = 0, phi = 0;

    for (;;) {
        const uint64_t dead = __ballot(!alive);
        if (dead != 0 && !exhausted) {
            const uint32_t ndead = (uint32_t)__popcll(dead);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(dead >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)dead, 0u));
            const uint64_t avail = pool_end - pool_next;
            uint64_t nb = 0, ne = 0;
            if (avail < ndead) {
                unsigned long long base = 0;
                if (lane == 0)
                    base = atomicAdd(a.counters, (unsigned long long)a.chunk);
                base = rfl64(base);
                if (base < a.photons) {
                    nb = base;
                    ne = base + a.chunk < a.photons ? base + a.chunk : a.photons;
                } else {
                    exhausted = true;
                }
            }
            if (!alive) {
                bool got = false;
                uint64_t idx = 0;
                if (rank < avail) {
                    idx = pool_next + rank;
                    got = true;
                } else if (rank - avail < ne - nb) {
                    idx = nb + (rank - avail);
                    got = true;
                }
                if (got) {
                    const uint64_t photon = a.first_photon + idx;
                    plo = (uint32_t)photon;
                    phi = (uint32_t)(photon >> 32);
                    x = y = z = 0.0f;
                    u = v = 0.0f;
                    w = 1.0f;
                    weight = 1.0f;
                    k = 0;
                    alive = true;
                }
            }
            if (avail >= ndead) {
                pool_next += ndead;
            } else {
                const uint64_t take = ndead - avail;
                pool_next = nb + (take < ne - nb ? take : ne - nb);
                pool_end = ne;
            }
        }
        const uint64_t live = __ballot(alive);
        if (live == 0)
            break;
        wave_interactions += (uint64_t)__popcll(live);

        bool truncated = false;
        if (alive) {
            const tmc_u32x4 r = tmc_philox4x32_10(k, 0u, plo, phi, key0, key1);

            /* move (photon.c:21-24) */
            const float t = -log32(tmc_u01(r.x0));
            x = __builtin_fmaf(t, u, x);
            y = __builtin_fmaf(t, v, y);
            z = __builtin_fmaf(t, w, z);

            /* absorb (photon.c:26-32) */
            const float r2 = __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x));
            const float f = sqrt_rn(r2) * a.spm;
            const uint32_t shell = f < a.shell_max ? (uint32_t)f : a.shells - 1;
            atomicAdd(&lw[shell], (unsigned long long)fx_heat(weight));
            atomicAdd(&lw2[shell], (unsigned long long)fx_heat2(weight));
            weight = weight * a.albedo;

            /* new direction (photon.c:34-43) */
            if (SAMPLER == 1u) {
                uint32_t w0 = r.x1, w1 = r.x2;
                tmc_u32x4 blk = { 0u, 0u, 0u, 0u };
                float xi1, xi2, tt;
                uint32_t att = 0;
                for (;;) {
                    xi1 = __builtin_fmaf(2.0f, tmc_u01(w0), -1.0f);
                    xi2 = __builtin_fmaf(2.0f, tmc_u01(w1), -1.0f);
                    tt = __builtin_fmaf(xi2, xi2, xi1 * xi1);
                    ++att;
                    if ((tt <= 1.0f && tt > 0.0f) || att >= 64u)
                        break;
                    if (att & 1u) {
                        blk = tmc_philox4x32_10(k, (att + 1u) >> 1, plo, phi, key0, key1);
                        w0 = blk.x0;
                        w1 = blk.x1;
                    } else {
                        w0 = blk.x2;
                        w1 = blk.x3;
                    }
                }
                u = __builtin_fmaf(2.0f, tt, -1.0f);
                const float s = 2.0f * sqrt_rn(1.0f - tt);
                v = xi1 * s;
                w = xi2 * s;
            } else {
                const float mu = __builtin_fmaf(2.0f, tmc_u01(r.x1), -1.0f);
                const float s = sqrt_rn(__builtin_fmaf(-mu, mu, 1.0f));
                float sn, cs;
                sincos2pi32(tmc_u01(r.x2), &sn, &cs);
                u = mu;
                v = s * cs;
                w = s * sn;
            }
            ++k;

            /* roulette (photon.c:45-49) */
            if (weight < 0.001f) {
                if (tmc_u01(r.x3) > 0.1f)
                    alive = false;
                else
                    weight = weight / 0.1f;
            }
            if (alive && k >= a.kcap) {
                alive = false;
                truncated = true;
            }
        }
        wave_truncated += (uint64_t)__popcll(__ballot(truncated));

        if (++since_flush >= a.flush_interval) {
            since_flush = 0;
            flush_tallies(lw, lw2, a.tally, a.shells, lane);
        }
    }
    flush_tallies(lw, lw2, a.tally, a.shells, lane);
    if (lane == 0) {
        atomicAdd(a.counters + 1, (unsigned long long)wave_interactions);
        if (wave_truncated)
            atomicAdd(a.counters + 2, (unsigned long long)wave_truncated);
    }
}

hipError_t launch_walk(uint32_t sampler, const WalkArgs& a, uint32_t grid, uint32_t block, hipStream_t stream)
{
    const size_t lds = (size_t)a.shells * 16u;
    if (sampler == 1u)
        hipLaunchKernelGGL(walk_f32<1u>, dim3(grid), dim3(block), lds, stream, a);
    else
        hipLaunchKernelGGL(walk_f32<0u>, dim3(grid), dim3(block), lds, stream, a);
    return hipGetLastError();
}

hipError_t walk_occupancy(uint32_t sampler, uint32_t block, uint32_t shells, int* wg_per_cu)
{
    const size_t lds = (size_t)shells * 16u;
    if (sampler == 1u)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, walk_f32<1u>, (int)block, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, walk_f32<0u>, (int)block, lds);
}

__global__ void philox_kernel(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const tmc_u32x4 r
        = tmc_philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], key[2 * i], key[2 * i + 1]);
    out[4 * i] = r.x0;
    out[4 * i + 1] = r.x1;
    out[4 * i + 2] = r.x2;
    out[4 * i + 3] = r.x3;
}

__global__ void eval_math_kernel(const float* in, uint32_t n, float* lg, float* sn, float* cs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float a = in[i];
    lg[i] = log32(a);
    float s, c;
    sincos2pi32(a, &s, &c);
    sn[i] = s;
    cs[i] = c;
}

hipError_t launch_philox(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(philox_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, ctr, key, n, out);
    return hipGetLastError();
}

hipError_t launch_eval_math(const float* in, uint32_t n, float* lg, float* sn, float* cs, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(eval_math_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, in, n, lg, sn, cs);
    return hipGetLastError();
}

} // namespace tmc
