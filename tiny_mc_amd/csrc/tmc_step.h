/* One absorption event of the photon walk in FP32 (photon.c:21-49), shared
 * by the single-run and multi-set kernels.  The operation order here is the
 * specification the host mirror (oracle) reproduces bit for bit. */
#ifndef TMC_STEP_H
#define TMC_STEP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tmc_math.h"
#include "tmc_philox.h"

namespace tmc {

/* Lane state of one photon (photon.c:11-18 plus its stream position). */
struct Photon {
    float x, y, z, u, v, w, weight;
    uint32_t k; /* absorptions so far = Philox counter word 0 */
    uint32_t plo, phi; /* photon index = Philox counter words 2-3 */
    tmc_philox_photon pp; /* the photon's part of Philox round 1 */
    /* rejection sampler: direction attempts made for absorption k so far
     * (0: none; the absorption itself is done once this is nonzero), with
     * bit 31 the absorption's cutoff flag, and its roulette word */
    uint32_t att, x3;
};

/* Physics of the photon's parameter set. */
struct Physics {
    float albedo; /* mu_s / (mu_s + mu_a) in binary32 (photon.c:8) */
    float spm; /* shells per mean free path (photon.c:9) */
    float shell_max; /* (float)(shells - 1), the overflow shell */
    uint32_t key0, key1; /* Philox key = 64-bit seed */
    float clear_half; /* 0.5 - eps, set by set_derived */
    tmc_philox_keys keys; /* round keys of (key0, key1), set by set_derived */
};

/* Per-set constants derived on the device: the Philox round keys, and the
 * bound of shell_index for the set's shell_max: a shell coordinate computed
 * from the hardware square root is within 3 ulp of the exact one (below), so
 * it has the exact floor if it lies more than eps = 8 ulp from every integer,
 * with the ulp of the floats just below 2^e > shell_max + 2; at or beyond
 * shell_max + 1 it is in the overflow shell either way. */
template <bool PIN_KEYS = false>
__device__ __forceinline__ void set_derived(Physics& ph)
{
    ph.keys = tmc_philox_key_schedule(ph.key0, ph.key1);
    if (PIN_KEYS) {
        /* where the keys change inside the walk loop (a new parameter set),
         * the compiler recomputes each round key inside the step (18 scalar
         * additions per absorption); the empty asm keeps them in scalar
         * registers instead */
#pragma unroll
        for (int r = 1; r < 10; ++r) {
            ph.keys.k0[r] = __builtin_amdgcn_readfirstlane(ph.keys.k0[r]);
            ph.keys.k1[r] = __builtin_amdgcn_readfirstlane(ph.keys.k1[r]);
            asm volatile("" : "+s"(ph.keys.k0[r]), "+s"(ph.keys.k1[r]));
        }
    }
    const int e = __builtin_amdgcn_frexp_expf(ph.shell_max + 2.0f); /* shell_max + 2 < 2^e */
    /* wave-uniform: keep the bound in a scalar register */
    ph.clear_half =
        __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(0.5f - __builtin_ldexpf(1.0f, e - 21))));
}

/* LDS words one lane deposits into.  Each pair of adjacent words holds a
 * shell's two sums (w image, w*w image), so one address serves both atomics
 * (the second takes the instruction's offset field).  Pairs 0 .. shell_max - 1
 * are the inner shells; the overflow shell (shell_max, the last one) has H
 * copies, pairs shell_max .. shell_max + H - 1: a long walk ends up there
 * most of the time, where the lanes of a wave would all hit one pair.  A lane
 * deposits at pair floor(min(f, fmax)), f its shell coordinate and
 * fmax = shell_max + 1/2 + (its copy, lane mod H): the clamp that bounds the
 * index also spreads the overflow deposits over the copies, which are added
 * up when the sums are flushed. */
struct LaneTally {
    unsigned long long* base; /* pair q at base[2 q], base[2 q + 1] */
    float fmax; /* shell_max + 0.5 + this lane's overflow copy */
    /* wave-uniform: the lanes whose deposits count (a walk may step lanes
     * that hold no photon, so that it never branches around the step) */
    uint64_t deposit;
};

/* Overflow copies tiny_mc_gpu_check_shell_index covers (copies 0 .. 31):
 * at least as many as either walk kernel uses. */
#define TMC_CHECK_HOT 32u

/* fmax of a lane for copy `copy` of the overflow shell. */
__device__ __forceinline__ float tally_fmax(float shell_max, uint32_t copy) { return shell_max + 0.5f + (float)copy; }

/* Adds one deposit's images to pair q: MASKED, for the lanes in lt.deposit
 * only; else for every active lane (a walk that steps only the lanes holding
 * a photon: there the mask would only split the step into blocks the
 * compiler schedules apart). */
template <bool MASKED = true>
__device__ __forceinline__ void tally_add(const LaneTally& lt, uint32_t q, unsigned long long a, unsigned long long b)
{
    if (!MASKED || __builtin_amdgcn_inverse_ballot_w64(lt.deposit)) {
        unsigned long long* p = lt.base + 2u * q;
        atomicAdd(p, a);
        atomicAdd(p + 1, b);
    }
}

/* The pair of shell s (inner) or of this lane's overflow copy, for a shell
 * coordinate computed exactly. */
__device__ __forceinline__ uint32_t tally_pair(const LaneTally& lt, bool inner, uint32_t s)
{
    return inner ? s : (uint32_t)lt.fmax;
}

/* The clamped short shell coordinate fc = min(RN(v_sqrt(r2) * spm), fmax)
 * and whether its floor is certainly the pair of the exact coordinate. */
__device__ __forceinline__ float shell_coord_short(float r2, const Physics& ph, float fmax, bool* sure)
{
    /* both operands are non-negative floats, ordered as their bit patterns:
     * one integer min, where fminf would first quiet a signalling fmax */
    const float fc = __uint_as_float(
        __builtin_elementwise_min(__float_as_uint(__builtin_amdgcn_sqrtf(r2) * ph.spm), __float_as_uint(fmax)));
    *sure = __builtin_fabsf(__builtin_amdgcn_fractf(fc) - 0.5f) < ph.clear_half;
    return fc;
}

/* Tally pair of a deposit at squared radius r2 (photon.c:26-29): with
 * f = RN(RN(sqrt(r2)) * spm), floor(f) if f < shell_max, else one of the
 * overflow shell's copies.  The hardware square root is within one ulp of
 * the correctly rounded one (the one-ulp correction of sqrt_cr_walk is exact
 * on every float), so f' = RN(v_sqrt(r2) * spm) is within 3 ulp of f.  The
 * clamped fc = min(f', fmax) has the right pair when it is clear of every
 * integer by eps: below shell_max + 1, floor(f') = floor(f) and both lie on
 * the same side of shell_max; from shell_max + 1 on (fmax included, whose
 * fraction is 1/2), f' and f are both overflows and any copy will do
 * (checked on every float r2 and every copy by tiny_mc_gpu_check_shell_index).
 * Otherwise, for a lane near a shell boundary (a few per ten thousand at 101
 * shells, including radii within eps / spm of the origin), the whole wave
 * takes the IEEE square root. */
__device__ __forceinline__ uint32_t shell_index(float r2, const Physics& ph, float fmax)
{
    bool sure;
    float fc = shell_coord_short(r2, ph, fmax, &sure);
    if (__builtin_expect(__ballot(!sure) != 0, 0))
        fc = sure ? fc : __uint_as_float(__builtin_elementwise_min(__float_as_uint(sqrt_rn(r2) * ph.spm),
                                                                    __float_as_uint(fmax)));
    return (uint32_t)fc;
}

__device__ __forceinline__ void launch_photon(Photon& p, uint64_t photon, uint32_t key0, uint32_t key1)
{
    p.x = p.y = p.z = 0.0f;
    p.u = p.v = 0.0f;
    p.w = 1.0f;
    p.weight = 1.0f;
    p.k = 0;
    p.att = 0;
    p.x3 = 0;
    p.plo = (uint32_t)photon;
    p.phi = (uint32_t)(photon >> 32);
    p.pp = tmc_philox_photon_init(p.plo, p.phi, key0, key1);
}

/* The binary32 weight sequence of one parameter set, as a table in LDS.
 * A photon's weight is a function of its absorption count k alone: every
 * photon starts at 1, multiplies by the albedo at each absorption
 * (photon.c:32) and, once below the cutoff, either dies or continues at the
 * weight divided by 0.1 (photon.c:45-49), the same for every survivor.
 * Entry k holds the fixed-point images of the weight W[k] deposited at
 * absorption k (fx_heat, fx_heat2: words 0-1 and 2-3, low word first).
 * Whether RN(w a) falls below the cutoff, for any weight w in [0, 1], is
 * fx_heat(w) < h_cut = fx_heat(T), T the least float with RN(T a) >= 0.001f:
 * RN(w a) is monotone in w, so the weights that fall below are those under
 * T, and T > 2^-10 (a < 1); fx_heat is exact, hence strictly increasing, on
 * [2^-10, 1], and below 2^23 = its image of 2^-10 underneath.  Photons at
 * k >= k_end (rare: the table spans many roulette cycles) fall back to the
 * arithmetic, starting from w_end = W[k_end]. */
struct WeightTable {
    const uint4* entry;
    uint32_t k_end;
    float w_end;
    uint64_t h_cut;
    /* wave-uniform: some photon of the wave may reach k_end before the walk
     * evaluates this again (table_near_end); while false, the step skips
     * the test of its photons against k_end */
    bool near_end;
};

/* Steps between evaluations of WeightTable::near_end. */
#define TMC_TABLE_CHECK_STEPS 64u

/* Whether a live photon of the wave (live: this lane holds one) can reach
 * k_end within the next TMC_TABLE_CHECK_STEPS steps: its k grows by at most
 * one per step, and photons launched meanwhile start at k = 0. */
template <class PhotonT>
__device__ __forceinline__ bool table_near_end(const PhotonT& ph, bool live, uint32_t k_end)
{
    if (k_end <= TMC_TABLE_CHECK_STEPS)
        return true;
    return __ballot(live && ph.k >= k_end - TMC_TABLE_CHECK_STEPS) != 0;
}

/* Whether a live photon of the wave (live: this lane holds one) can reach
 * the absorption cap kcap within the next `steps` steps: its k grows by one
 * per step, and photons
 * launched meanwhile start at k = 0.  Evaluated at each periodic flush, it
 * lets the walk skip the per-step cap check until the next one; with the
 * default cap (2^32 - 1) the check never runs. */
template <class PhotonT>
__device__ __forceinline__ bool cap_within(const PhotonT& ph, bool live, uint32_t kcap, uint32_t steps)
{
    if (kcap <= steps)
        return true;
    return __ballot(live && ph.k >= kcap - steps) != 0;
}

/* The roulette kills a photon when u > 0.1f (photon.c:46).  For
 * u = ((r >> 8) + 1) 2^-24 that is (r >> 8) + 1 > 0.1f * 2^24 = 1677721.625,
 * i.e. r >= 1677721 * 2^8: one integer compare on the Philox word. */
#define TMC_ROULETTE_KILL_R 0x19999900u

/* Whether the walk's next iteration of p is an absorption (the rejection
 * sampler's retries of a direction are iterations without one). */
template <uint32_t SAMPLER>
__device__ __forceinline__ bool absorbs(const Photon& p)
{
    return SAMPLER == 0u || p.att == 0u;
}

/* Walks one iteration of p: with the analytic sampler, and with the
 * rejection sampler when the first attempt is accepted, one absorption
 * event, its deposit added to the fixed-point shell sums in LDS.  Returns
 * false when the roulette kills the photon.
 * MATH (TMC_MATH_*) selects how -ln u, sin/cos(2 pi u') and the analytic
 * sampler's sin(theta) are evaluated; everything else is the same
 * arithmetic in both modes.  The Philox round keys come from ph.keys
 * (set_derived).  TABLE reads the weight's images from wt (WeightTable).
 * Rejection sampler (photon.c:34-43): attempt 1 uses words 1-2 of block
 * (k, 0), attempts 2 and 3 words 0-1 and 2-3 of block (k, 1), and so on.
 * An iteration makes at most the attempts of one block: a photon whose
 * attempt is rejected keeps its absorption pending (position, deposit and
 * weight done, direction, k and roulette not) and makes the next attempts
 * in the next iteration, from the next block.  Its lanes thus compute one
 * block per iteration like every other lane, instead of the whole wave
 * computing a second block whenever one of its 64 lanes rejects (almost
 * every time: 1 - (pi/4)^64).  The arithmetic and its order per absorption
 * are unchanged. */
/* What the roulette of a step reads: whether the weight after this
 * absorption fell below the cutoff (false while a rejection attempt is
 * pending) and the roulette word. */
struct StepEnd {
    bool cut;
    uint32_t x3;
    /* with the table and the analytic sampler: the wave's ballot of cut,
     * taken before the deposit (below) */
    uint64_t cut_lanes;
};

template <uint32_t SAMPLER, uint32_t MATH, bool TABLE, bool MASKED, bool R2 = false>
__device__ __forceinline__ StepEnd step_core(Photon& p, const Physics& ph, const LaneTally& lt, const WeightTable& wt)
{
    const bool fresh = absorbs<SAMPLER>(p);
    /* the table entry first: its load overlaps the generator (where exactly
     * is left to the compiler: pinning the load ahead of the generator with
     * a scheduling barrier measured 1.3% slower) */
    uint4 e = { 0u, 0u, 0u, 0u };
    if (TABLE)
        e = wt.entry[__builtin_elementwise_min(p.k, wt.k_end - 1u)];
    tmc_u32x4 r;
    if (SAMPLER == 1u) /* block (k, 0) for an absorption, (k, b) for attempts 2b and 2b + 1 */
        r = tmc_philox4x32_10_blockb(p.k, (p.att & 0xFFu) ? ((p.att & 0xFFu) + 1u) >> 1 : 0u, p.pp, &ph.keys,
                                     ph.key0, ph.key1);
    else if (R2) /* block (k, 0), round 2's photon-constant product from pp */
        r = tmc_philox4x32_10_block0_r2(p.k, p.pp, ph.keys);
    else /* block (k, 0) */
        r = tmc_philox4x32_10_block0_ks(p.k, p.pp, ph.keys);

    bool cut = false; /* the weight after this absorption is below the cutoff */
    uint64_t cut_lanes = 0;
    if (fresh) {
        /* move (photon.c:21-24) */
        const float t = step_length<MATH>(r.x0);
        p.x = __builtin_fmaf(t, p.u, p.x);
        p.y = __builtin_fmaf(t, p.v, p.y);
        p.z = __builtin_fmaf(t, p.w, p.z);

        /* absorb (photon.c:26-32) */
        const float r2 = __builtin_fmaf(p.z, p.z, __builtin_fmaf(p.y, p.y, p.x * p.x));
        const uint32_t q = shell_index(r2, ph, lt.fmax);
        if (TABLE) {
            uint64_t h = ((uint64_t)e.y << 32) | e.x, h2 = ((uint64_t)e.w << 32) | e.z;
            if (__builtin_expect(wt.near_end, 0) && __builtin_expect(__ballot(p.k >= wt.k_end) != 0, 0)) {
                /* photons past the table continue with the arithmetic */
                if (p.k >= wt.k_end) {
                    if (p.k == wt.k_end)
                        p.weight = wt.w_end;
                    h = fx_heat(p.weight);
                    h2 = fx_heat2(p.weight);
                    p.weight = p.weight * ph.albedo;
                    /* the weight a roulette survivor continues with (a
                     * photon the roulette kills drops it) */
                    if (p.weight < 0.001f)
                        p.weight = div_tenth(p.weight); /* = p.weight / 0.1f */
                }
            }
            cut = h < wt.h_cut; /* = RN(w a) < 0.001f for the deposited weight w */
            /* the ballot here, ahead of the masked deposit: with the compare
             * left for after it, the compiler waited at the join for the
             * table read and, with it, for both LDS atomics (its counter
             * is in order) */
            if (SAMPLER == 0u && MASKED)
                cut_lanes = __ballot(cut);
            tally_add<MASKED>(lt, q, h, h2);
        } else {
            tally_add<MASKED>(lt, q, fx_heat(p.weight), fx_heat2(p.weight));
            p.weight = p.weight * ph.albedo;
            cut = p.weight < 0.001f;
        }
    }

    /* new direction (photon.c:34-43) */
    uint32_t x3 = r.x3; /* the roulette word of block (k, 0) */
    if (SAMPLER == 1u) {
        uint32_t att = p.att & 0xFFu;
        float xi1 = tmc_u11(fresh ? r.x1 : r.x0), xi2 = tmc_u11(fresh ? r.x2 : r.x1);
        float tt = __builtin_fmaf(xi2, xi2, xi1 * xi1);
        ++att;
        bool accept = (tt <= 1.0f && tt > 0.0f) || att >= 64u;
        if (!accept && !fresh) {
            xi1 = tmc_u11(r.x2);
            xi2 = tmc_u11(r.x3);
            tt = __builtin_fmaf(xi2, xi2, xi1 * xi1);
            ++att;
            accept = (tt <= 1.0f && tt > 0.0f) || att >= 64u;
        }
        if (!accept) {
            /* pending: the next iteration makes the next attempts */
            if (fresh) {
                p.x3 = r.x3;
                p.att = att | (cut ? 0x80000000u : 0u);
            } else {
                p.att = (p.att & 0x80000000u) | att;
            }
            return StepEnd{ false, 0u, 0ull };
        }
        if (!fresh) {
            cut = (p.att & 0x80000000u) != 0u;
            x3 = p.x3;
        }
        p.att = 0;
        p.u = __builtin_fmaf(2.0f, tt, -1.0f);
        const float s = 2.0f * sqrt_cr_walk(1.0f - tt); /* 1 - tt: 0 or >= 2^-24 */
        p.v = xi1 * s;
        p.w = xi2 * s;
    } else {
        const float mu = tmc_u11(r.x1); /* = 2 u - 1, exact */
        const float s = sin_polar<MATH>(mu);
        float sn, cs;
        sincos_turn<MATH>(r.x2, &sn, &cs);
        p.u = mu;
        p.v = s * cs;
        p.w = s * sn;
    }
    ++p.k;
    return StepEnd{ cut, x3, cut_lanes };
}

/* One iteration of p (step_core), then the roulette (photon.c:45-49).
 * Returns false when the roulette kills the photon.  MASKED: deposits for
 * the lanes in lt.deposit only (tally_add).  R2: the analytic sampler's
 * block takes Philox round 2's photon-constant product from the photon
 * (tmc_philox4x32_10_block0_r2): one multiplication fewer per step, two
 * more registers per photon; measured on MI355X 1.8% faster in the batch
 * walk, 1.0% slower in the single-run walk without the weight table and
 * 0.3% slower with it (same bits in all three). */
template <uint32_t SAMPLER, uint32_t MATH, bool TABLE = false, bool MASKED = false, bool R2 = false>
__device__ __forceinline__ bool step(Photon& p, const Physics& ph, const LaneTally& lt,
                                     const WeightTable& wt = WeightTable{ nullptr, 0u, 0.0f, 0ull, false })
{
    const StepEnd e = step_core<SAMPLER, MATH, TABLE, MASKED, R2>(p, ph, lt, wt);
    if (e.cut) {
        if (e.x3 >= TMC_ROULETTE_KILL_R) /* u > 0.1f */
            return false;
        if (!TABLE) /* (with the table, a photon past it has divided already) */
            p.weight = div_tenth(p.weight); /* = p.weight / 0.1f */
    }
    return true;
}

/* step() for a whole wave whose lanes all step (those without a photon
 * with their deposits masked): returns the lanes whose photon the roulette
 * kills, as the ballots of the two compares it reads.  (The ballot of
 * step()'s result costs two vector instructions: the compiler rebuilds a
 * compare from the merged lane mask.) */
template <uint32_t SAMPLER, uint32_t MATH, bool TABLE>
__device__ __forceinline__ uint64_t step_wave(Photon& p, const Physics& ph, const LaneTally& lt, const WeightTable& wt)
{
    const StepEnd e = step_core<SAMPLER, MATH, TABLE, true>(p, ph, lt, wt);
    /* (taken before the branch below, across which the compiler would
     * rebuild the cut flag's ballot from a merged lane mask) */
    const uint64_t cut_lanes = TABLE && SAMPLER == 0u ? e.cut_lanes : __ballot(e.cut);
    const uint64_t ends = cut_lanes & __ballot(e.x3 >= TMC_ROULETTE_KILL_R); /* u > 0.1f */
    if (!TABLE && e.cut) /* (a photon the roulette kills drops its weight) */
        p.weight = div_tenth(p.weight); /* = p.weight / 0.1f */
    return ends;
}

} // namespace tmc

#endif
