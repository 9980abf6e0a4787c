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
#include "tmc_device.h"
#include "tmc_internal.h"
#include "tmc_math.h"
#include "tmc_philox.h"
#include "tmc_step.h"
#include "tmc_step_prec.h"

namespace tmc {

/* Copies of the overflow shell's sums per workgroup (lane & 7 picks one's
 * bound, LaneTally).  Measured per launch (2^30 photons at default physics,
 * 2^22 at mu_a = 0.01, mu_s = 100 over 401 x 10 um, 2^26 at albedo 0.98 over
 * 201 x 20 um): 32 copies 136.7, 654, 46.2 ms; 16: 134.6, 657, 44.2; 8:
 * 134.0, 656, 44.1; 4: 135.6, 654, 44.1; 2: 136.1, 1077, 66.5. */
#ifndef TMC_WALK_HOT
#define TMC_WALK_HOT 8u
#endif
static_assert(TMC_WALK_HOT <= TMC_CHECK_HOT, "the shell check covers every overflow copy");

/* Dead lanes a wave gathers before it launches photons into them: a launch
 * costs the wave about fifteen vector instructions whatever the number of
 * lanes it fills, and the lanes wait a few steps in exchange.  Measured at
 * 2^30 photons, default physics: 1 -> 149.9 ms, 2 -> 144.4, 4 -> 143.6,
 * 8 -> 143.6. */
#ifndef TMC_WALK_RELAUNCH_MIN
#define TMC_WALK_RELAUNCH_MIN 4
#endif
/* a wave whose lanes are all dead must launch, or it would leave the walk */
static_assert(TMC_WALK_RELAUNCH_MIN >= 1 && TMC_WALK_RELAUNCH_MIN <= 64, "a drained wave must relaunch");

/* LDS words of one workgroup: the (w, w*w) sum pairs of the inner shells,
 * then the overflow shell's copies (LaneTally; a multiple of 2 words: the
 * table that follows is 16-byte aligned). */
__host__ __device__ __forceinline__ size_t walk_lds_words(uint32_t shells)
{
    return 2u * ((size_t)shells - 1u + TMC_WALK_HOT);
}

/* LDS bytes of a workgroup: the tallies, then table_k table entries (16
 * bytes) and the weight sequence they are built from (table_k + 1 floats). */
__host__ __device__ __forceinline__ size_t walk_lds_bytes(uint32_t shells, uint32_t table_k)
{
    return walk_lds_words(shells) * sizeof(unsigned long long)
        + (table_k ? 16u * (size_t)table_k + 4u * ((size_t)table_k + 3u) : 0u);
}

/* Builds the weight table of a binary32 walk in LDS (WeightTable): one
 * thread runs the recurrence W[0] = 1, W[k+1] = RN(W[k] a), divided by 0.1
 * when it falls below the cutoff (the arithmetic of step()), and takes the
 * cut threshold h_cut, then every thread converts its entries.  Ends with a
 * barrier. */
__device__ __forceinline__ WeightTable build_weight_table(unsigned long long* lds, uint32_t shells, uint32_t K,
                                                          float albedo)
{
    uint4* tab = (uint4*)(lds + walk_lds_words(shells));
    float* wseq = (float*)(tab + K); /* W[0 .. K], then h_cut as two words */
    uint32_t* hcut = (uint32_t*)(wseq + K + 1);
    if (threadIdx.x == 0) {
        float w = 1.0f;
        for (uint32_t k = 0; k < K; ++k) {
            wseq[k] = w;
            const float w2 = w * albedo;
            w = w2 < 0.001f ? div_tenth(w2) : w2;
        }
        wseq[K] = w;
        /* T: the least float with RN(T a) >= 0.001f, by bisection over the
         * bit patterns of (0, 1] (ordered as the floats); none when even
         * 1 * a falls below, and then every weight does */
        uint64_t h_cut = ~0ull;
        if (!(1.0f * albedo < 0.001f)) {
            uint32_t lo = 0u, hi = __float_as_uint(1.0f); /* RN(lo a) < 0.001f <= RN(hi a) */
            while (hi - lo > 1u) {
                const uint32_t mid = lo + (hi - lo) / 2u;
                if (__uint_as_float(mid) * albedo < 0.001f)
                    lo = mid;
                else
                    hi = mid;
            }
            h_cut = fx_heat(__uint_as_float(hi));
        }
        hcut[0] = (uint32_t)h_cut;
        hcut[1] = (uint32_t)(h_cut >> 32);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) {
        const float w = wseq[k];
        const uint64_t h = fx_heat(w), h2 = fx_heat2(w);
        tab[k] = make_uint4((uint32_t)h, (uint32_t)(h >> 32), (uint32_t)h2, (uint32_t)(h2 >> 32));
    }
    __syncthreads();
    const uint64_t h_cut = ((uint64_t)hcut[1] << 32) | hcut[0];
    return WeightTable{ tab, K, wseq[K], h_cut, true };
}

/* Moves this workgroup's LDS tallies (inner shell pairs, then the overflow
 * shell's copies, each a (w image, w*w image) pair of words) into the global
 * 128-bit sums.  Safe while other waves of the workgroup keep adding: each
 * entry is taken by an atomic swap with zero, so every addition is counted
 * exactly once. */
__device__ __forceinline__ void flush_tallies(unsigned long long* lds, unsigned long long* g, uint32_t shells,
                                              uint32_t lane)
{
    const uint32_t last = shells - 1;
    for (uint32_t q = lane; q < last + TMC_WALK_HOT; q += 64) {
        const unsigned long long a = atomicExch(&lds[2 * q], 0ull);
        const unsigned long long b = atomicExch(&lds[2 * q + 1], 0ull);
        const uint32_t s = q < last ? q : last;
        if (a)
            add128(g + 2 * s, a);
        if (b)
            add128(g + 2 * shells + 2 * s, b);
    }
}

/* One iteration of every lane of the wave, by photon state type: the
 * reference's FP32 step or one of the precision variants
 * (tmc_step_prec.h); returns the lanes whose photon it ends.
 * VARIANT: the math mode for the binary32 walk, compensation for the
 * weight-from-counts walk, unused for binary64. */
template <uint32_t SAMPLER, uint32_t VARIANT, bool TABLE>
__device__ __forceinline__ uint64_t step_of(Photon& p, const Physics& ph, const LaneTally& lt, const WeightTable& wt)
{
    if (SAMPLER == 0u)
        return step_wave<SAMPLER, VARIANT, TABLE>(p, ph, lt, wt);
    /* (the rejection sampler's pending attempts merge its cut flag across a
     * branch: step_wave measured 1.5% slower for it) */
    return __ballot(!step<SAMPLER, VARIANT, TABLE, true>(p, ph, lt, wt));
}

template <uint32_t SAMPLER, uint32_t VARIANT, bool TABLE>
__device__ __forceinline__ uint64_t step_of(PhotonW& p, const PhysicsW& ph, const LaneTally& lt, const WeightTable&)
{
    return __ballot(!step<SAMPLER, VARIANT != 0u>(p, ph, lt));
}

template <uint32_t SAMPLER, uint32_t VARIANT, bool TABLE>
__device__ __forceinline__ uint64_t step_of(Photon64& p, const Physics64& ph, const LaneTally& lt, const WeightTable&)
{
    return __ballot(!step<SAMPLER>(p, ph, lt));
}

/* Whether the next iteration of a live photon is an absorption (only the
 * binary32 walk's rejection sampler has iterations that are not). */
template <uint32_t SAMPLER>
__device__ __forceinline__ bool absorbs_of(const Photon& p)
{
    return absorbs<SAMPLER>(p);
}
template <uint32_t SAMPLER>
__device__ __forceinline__ bool absorbs_of(const PhotonW&)
{
    return true;
}
template <uint32_t SAMPLER>
__device__ __forceinline__ bool absorbs_of(const Photon64&)
{
    return true;
}

/* The persistent walk over photons [first_photon, first_photon + photons). */
template <class PhotonT, class PhysicsT, uint32_t SAMPLER, uint32_t VARIANT, bool TABLE = false>
__device__ __forceinline__ void walk_body(const WalkArgs& a, const PhysicsT& phys)
{
    extern __shared__ unsigned long long lds_tally[];
    for (uint32_t i = threadIdx.x; i < walk_lds_words(a.shells); i += blockDim.x)
        lds_tally[i] = 0;
    WeightTable wt = { nullptr, 0u, 0.0f, 0ull, false };
    if (TABLE)
        wt = build_weight_table(lds_tally, a.shells, a.table_k, a.albedo); /* ends with a barrier */
    else
        __syncthreads();

    const uint32_t lane = __lane_id();
    LaneTally lt = { lds_tally, tally_fmax(a.shell_max, lane & (TMC_WALK_HOT - 1)), 0ull };

    /* wave-uniform pool of claimed photon offsets [pool_next, pool_end) */
    uint64_t pool_next = 0, pool_end = 0;
    bool exhausted = false;
    uint64_t wave_interactions = 0, wave_truncated = 0;
    uint32_t since_flush = 0;

    /* lane state: one photon (photon.c:11-18); the lanes without one are
     * the wave-uniform mask `dead` (scalar registers: a lane's liveness
     * costs no vector compare, and the step runs under the mask itself) */
    PhotonT ph;
    launch_photon(ph, 0, phys.key0, phys.key1);
    uint64_t dead = __builtin_amdgcn_read_exec(); /* none has one yet */
    bool capping = cap_within(ph, false, a.kcap, a.flush_interval); /* per-step cap checks until the next flush */
    if (TABLE)
        wt.near_end = table_near_end(ph, false, wt.k_end);

    for (;;) {
        /* the dead lanes take photons from the pool once there are
         * TMC_WALK_RELAUNCH_MIN of them (at most 64, so a drained wave always
         * does); an empty pool is refilled with the next chunk first.  Lanes the pool cannot serve
         * wait for a later claim, so every launch indexes one contiguous
         * range (pool_next + rank). */
        /* (a 32-bit count: compared as 64 bits it takes a vector compare) */
        if ((uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)__popcll(dead)) >= TMC_WALK_RELAUNCH_MIN
            && !exhausted) {
            if (pool_next == pool_end) {
                unsigned long long base = 0;
                if (lane == 0)
                    base = atomicAdd(a.counters, (unsigned long long)a.chunk);
                base = rfl64(base);
                if (base < a.photons) {
                    pool_next = base;
                    pool_end = base + a.chunk < a.photons ? base + a.chunk : a.photons;
                } else {
                    exhausted = true;
                }
            }
            if (!exhausted) {
                const uint32_t n = (uint32_t)__popcll(dead);
                const uint32_t take = pool_take(pool_next, pool_end, n);
                const uint64_t first = a.first_photon + pool_next;
                const bool gets = __builtin_amdgcn_inverse_ballot_w64(dead) && mask_rank(dead) < take;
                if (gets)
                    launch_photon(ph, first + mask_rank(dead), phys.key0, phys.key1);
                /* the usual case, every dead lane served, without a ballot */
                dead = take == n ? 0ull : dead & ~__ballot(gets);
                pool_next += take;
            }
        }
        const uint64_t live = __builtin_amdgcn_read_exec() & ~dead;
        if (live == 0)
            break;
        const bool lane_live = __builtin_amdgcn_inverse_ballot_w64(live);
        wave_interactions +=
            (uint64_t)__popcll(SAMPLER == 1u ? __ballot(lane_live && absorbs_of<SAMPLER>(ph)) : live);

        /* every lane steps, its deposits masked by `live`: a lane without a
         * photon walks a stale one, which a launch overwrites.  Around a
         * branch on liveness, the lanes the step ends would be a lane mask
         * merged at the join, whose ballot costs two vector instructions per
         * step; here it is the step's own compare. */
        lt.deposit = live;
        dead |= step_of<SAMPLER, VARIANT, TABLE>(ph, phys, lt, wt) & live;
        /* photons that reached the absorption cap stop; checked only while
         * some photon of the wave can reach it before the next flush */
        if (capping) {
            const uint64_t capped = __ballot(ph.k >= a.kcap) & live & ~dead;
            if (capped != 0) {
                wave_truncated += (uint64_t)__popcll(capped);
                dead |= capped;
            }
        }

        /* wave-uniform; readfirstlane keeps it in a scalar register */
        since_flush = __builtin_amdgcn_readfirstlane(since_flush + 1u);
        if (since_flush >= a.flush_interval) {
            since_flush = 0;
            flush_tallies(lds_tally, a.tally, a.shells, lane);
            capping = cap_within(ph, !__builtin_amdgcn_inverse_ballot_w64(dead), a.kcap, a.flush_interval);
        }
        /* at most TMC_TABLE_CHECK_STEPS steps apart, flushes included */
        if (TABLE && since_flush % TMC_TABLE_CHECK_STEPS == 0u)
            wt.near_end = table_near_end(ph, !__builtin_amdgcn_inverse_ballot_w64(dead), wt.k_end);
    }
    flush_tallies(lds_tally, a.tally, a.shells, lane);
    if (lane == 0) {
        atomicAdd(a.counters + 1, (unsigned long long)wave_interactions);
        if (wave_truncated)
            atomicAdd(a.counters + 2, (unsigned long long)wave_truncated);
    }
}

/* The reference's FP32 arithmetic (the production walk); TABLE: the
 * weight sequence from the LDS table (WeightTable). */
template <uint32_t SAMPLER, uint32_t MATH, bool TABLE>
__global__ void __launch_bounds__(1024) walk_f32(WalkArgs a)
{
    Physics phys = {};
    phys.albedo = a.albedo;
    phys.spm = a.spm;
    phys.shell_max = a.shell_max;
    phys.key0 = (uint32_t)a.seed;
    phys.key1 = (uint32_t)(a.seed >> 32);
    set_derived(phys);
    walk_body<Photon, Physics, SAMPLER, MATH, TABLE>(a, phys);
}

/* FP32 transport, weight from exact counts (COMP: compensated positions). */
template <uint32_t SAMPLER, bool COMP>
__global__ void __launch_bounds__(1024) walk_f32w(WalkArgs a)
{
    const PhysicsW phys = { a.l2a_hi, a.l2a_lo, a.l2s_hi, a.l2s_lo, a.spm_w, a.shell_max, a.shells - 1,
                            (uint32_t)a.seed, (uint32_t)(a.seed >> 32) };
    walk_body<PhotonW, PhysicsW, SAMPLER, COMP ? 1u : 0u>(a, phys);
}

/* All-FP64 transport. */
template <uint32_t SAMPLER>
__global__ void __launch_bounds__(1024) walk_f64(WalkArgs a)
{
    const Physics64 phys = { a.albedo_d, a.spm_d, (double)(a.shells - 1), a.shells - 1, (uint32_t)a.seed,
                             (uint32_t)(a.seed >> 32) };
    walk_body<Photon64, Physics64, SAMPLER, 0u>(a, phys);
}

namespace {

using WalkKernel = void (*)(WalkArgs);

WalkKernel walk_kernel(uint32_t sampler, uint32_t precision, uint32_t math, bool table)
{
    const bool rej = sampler == 1u;
    switch (precision) {
    case 1u: return rej ? walk_f32w<1u, false> : walk_f32w<0u, false>;
    case 2u: return rej ? walk_f32w<1u, true> : walk_f32w<0u, true>;
    case 3u: return rej ? walk_f64<1u> : walk_f64<0u>;
    default:
        if (math == TMC_MATH_OWN) {
            if (table)
                return rej ? walk_f32<1u, TMC_MATH_OWN, true> : walk_f32<0u, TMC_MATH_OWN, true>;
            return rej ? walk_f32<1u, TMC_MATH_OWN, false> : walk_f32<0u, TMC_MATH_OWN, false>;
        }
        if (table)
            return rej ? walk_f32<1u, TMC_MATH_HARDWARE, true> : walk_f32<0u, TMC_MATH_HARDWARE, true>;
        return rej ? walk_f32<1u, TMC_MATH_HARDWARE, false> : walk_f32<0u, TMC_MATH_HARDWARE, false>;
    }
}

} // namespace

hipError_t launch_walk(uint32_t sampler, uint32_t precision, uint32_t math, const WalkArgs& a, uint32_t grid,
                       uint32_t block, hipStream_t stream)
{
    const uint32_t K = precision == 0u ? a.table_k : 0u;
    hipLaunchKernelGGL(walk_kernel(sampler, precision, math, K != 0u), dim3(grid), dim3(block),
                       walk_lds_bytes(a.shells, K), stream, a);
    return hipGetLastError();
}

hipError_t walk_occupancy(uint32_t sampler, uint32_t precision, uint32_t math, uint32_t block, uint32_t shells,
                          uint32_t table_k, int* wg_per_cu)
{
    const uint32_t K = precision == 0u ? table_k : 0u;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, walk_kernel(sampler, precision, math, K != 0u),
                                                        (int)block, walk_lds_bytes(shells, K));
}

__global__ void clear_kernel(unsigned long long* p, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = 0;
}

hipError_t launch_clear(unsigned long long* p, size_t n, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    const size_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(clear_kernel, dim3(blocks < 1024 ? (uint32_t)blocks : 1024u), dim3(256), 0, stream, p, n);
    return hipGetLastError();
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

/* The walk's short forms against the IEEE operation they replace. */
template <uint32_t OP>
__device__ __forceinline__ bool short_op_matches(float x)
{
    if (OP == TMC_CHECK_SQRT)
        return __float_as_uint(sqrt_cr_walk(x)) == __float_as_uint(sqrt_rn(x));
    if (OP == TMC_CHECK_DIV_TENTH_SHORT)
        return __float_as_uint(div_tenth_short(x)) == __float_as_uint(x / 0.1f);
    return __float_as_uint(div_tenth(x)) == __float_as_uint(x / 0.1f);
}

/* Compares a short form with the IEEE operation on the floats with bit
 * patterns [first, first + count): out[0] += mismatches, out[1] = min of the
 * mismatching patterns. */
template <uint32_t OP>
__global__ void check_short_op_kernel(uint32_t first, uint64_t count, unsigned long long* out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0, lowest = ~0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint32_t bits = first + (uint32_t)i;
        const float x = __uint_as_float(bits);
        if (!short_op_matches<OP>(x)) {
            ++bad;
            lowest = bits < lowest ? bits : lowest;
        }
    }
    if (bad) {
        atomicAdd(&out[0], bad);
        atomicMin(&out[1], lowest);
    }
}

/* Checks shell_index's claim on the squared radii r2 with bit patterns
 * [first, first + count) for one set's spm and shell_max, for each of the
 * overflow copies 0 .. TMC_CHECK_HOT - 1 (fmax = shell_max + 1/2 + copy):
 * where the clamped short coordinate fc is marked sure, floor(fc) must be
 * floor(fe) for fe = RN(RN(sqrt(r2)) * spm) < shell_max, and a pair from
 * shell_max to shell_max + copy otherwise.  out[0] += violations (over all
 * copies), out[1] = min of the violating patterns, out[2] += radii not marked
 * sure for copy 0. */
__global__ void check_shell_kernel(float spm, float shell_max, uint32_t first, uint64_t count,
                                   unsigned long long* out)
{
    Physics ph = {};
    ph.spm = spm;
    ph.shell_max = shell_max;
    set_derived(ph);
    const uint32_t last = (uint32_t)shell_max;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0, lowest = ~0ull, unsure = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint32_t bits = first + (uint32_t)i;
        const float r2 = __uint_as_float(bits);
        const float fe = sqrt_rn(r2) * spm;
        for (uint32_t copy = 0; copy < TMC_CHECK_HOT; ++copy) {
            bool sure;
            const float fc = shell_coord_short(r2, ph, tally_fmax(shell_max, copy), &sure);
            if (!sure) {
                unsure += copy == 0u;
                continue;
            }
            const uint32_t q = (uint32_t)fc;
            const bool ok = fe < shell_max ? q == (uint32_t)fe : q >= last && q <= last + copy;
            if (!ok) {
                ++bad;
                lowest = bits < lowest ? bits : lowest;
            }
        }
    }
    if (bad) {
        atomicAdd(&out[0], bad);
        atomicMin(&out[1], lowest);
    }
    if (unsure)
        atomicAdd(&out[2], unsure);
}

hipError_t launch_check_shell(float spm, float shell_max, uint32_t first, uint64_t count, unsigned long long* out,
                              hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(check_shell_kernel, dim3(8192), dim3(256), 0, stream, spm, shell_max, first, count, out);
    return hipGetLastError();
}

hipError_t launch_check_short_op(uint32_t op, uint32_t first, uint64_t count, unsigned long long* out,
                                 hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    if (op == TMC_CHECK_SQRT)
        hipLaunchKernelGGL(check_short_op_kernel<TMC_CHECK_SQRT>, dim3(8192), dim3(256), 0, stream, first, count,
                           out);
    else if (op == TMC_CHECK_DIV_TENTH_SHORT)
        hipLaunchKernelGGL(check_short_op_kernel<TMC_CHECK_DIV_TENTH_SHORT>, dim3(8192), dim3(256), 0, stream,
                           first, count, out);
    else
        hipLaunchKernelGGL(check_short_op_kernel<TMC_CHECK_DIV_TENTH>, dim3(8192), dim3(256), 0, stream, first,
                           count, out);
    return hipGetLastError();
}

/* The step length and sin/cos(2 pi u) of a binary32 walk in math mode MATH
 * on inputs u = in[i] (each of the form ((r >> 8) + 1) 2^-24). */
template <uint32_t MATH>
__global__ void eval_step_math_kernel(const float* in, uint32_t n, float* t, float* sn, float* cs, float* sp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    /* recover r >> 8 from u: u * 2^24 - 1 is an exact integer in [0, 2^24) */
    const uint32_t r = ((uint32_t)(in[i] * 0x1p24f) - 1u) << 8;
    t[i] = step_length<MATH>(r);
    float s, c;
    sincos_turn<MATH>(r, &s, &c);
    sn[i] = s;
    cs[i] = c;
    /* the analytic sampler's sin(theta) for mu = 2u - 1 */
    sp[i] = sin_polar<MATH>(tmc_u11(r));
}

hipError_t launch_eval_step_math(uint32_t math, const float* in, uint32_t n, float* t, float* sn, float* cs,
                                 float* sp, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    if (math == TMC_MATH_OWN)
        hipLaunchKernelGGL(eval_step_math_kernel<TMC_MATH_OWN>, dim3((n + 255) / 256), dim3(256), 0, stream, in, n,
                           t, sn, cs, sp);
    else
        hipLaunchKernelGGL(eval_step_math_kernel<TMC_MATH_HARDWARE>, dim3((n + 255) / 256), dim3(256), 0, stream,
                           in, n, t, sn, cs, sp);
    return hipGetLastError();
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
