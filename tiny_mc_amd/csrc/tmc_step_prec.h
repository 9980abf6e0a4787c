/* Precision variants of one absorption event (photon.c:21-49), for the
 * high-albedo precision study.  The reference's FP32 step is tmc_step.h;
 * these walk the same Philox stream (same words for step, direction and
 * roulette) with other arithmetic:
 *
 *   StepWeightF32<SAMPLER, COMP>
 *     FP32 transport whose weight is not carried by repeated products
 *     (photon.c:32 `weight *= albedo`, which drifts by about -1.8e-3 by the
 *     cutoff at albedo 0.9999) but derived from exact integers: after e
 *     absorptions and r roulette survivals the weight is
 *         albedo^e * (1/0.1f)^r = 2^(e log2(albedo) - r log2(0.1f)),
 *     with log2(albedo) and log2(0.1f) computed in FP64 on the host and
 *     split into float pairs; the exponent is summed with error-free
 *     transformations and the one rounding of the hardware exp2 remains.
 *     The shells-per-mean-free-path factor is also computed in FP64 and
 *     rounded once.  COMP additionally keeps each coordinate as an
 *     unevaluated float pair (compensated summation of x += t u), so
 *     position error no longer grows with the float spacing of |x|.
 *
 *   StepF64<SAMPLER>
 *     Every quantity in FP64 (step, position, direction, weight, radius)
 *     with the device library's log / sqrt / sincos: the GPU counterpart of
 *     the FP64 oracle (same expressions, no contraction), so it agrees with
 *     that oracle up to the last-place differences of log and sincos.
 *
 * Both deposit into the same fixed-point images as the FP32 step
 * (w * 2^33 and w^2 * 2^43): exact for a float weight >= 2^-10, rounded to
 * nearest for an FP64 weight. */
#ifndef TMC_STEP_PREC_H
#define TMC_STEP_PREC_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tmc_math.h"
#include "tmc_philox.h"
#include "tmc_step.h"

namespace tmc {

/* ---------------------------------------------------------------------- */
/* FP32 with the weight derived from exact counts                          */
/* ---------------------------------------------------------------------- */

struct PhotonW {
    float x, y, z, u, v, w;
    float xl, yl, zl; /* low parts of the coordinates (COMP only) */
    float weight;
    uint32_t e; /* albedo factors applied */
    uint32_t r; /* roulette survivals */
    uint32_t k; /* absorptions so far = Philox counter word 0 */
    uint32_t plo, phi;
    tmc_philox_photon pp;
};

struct PhysicsW {
    float l2a_hi, l2a_lo; /* log2(albedo), albedo = mu_s / (mu_s + mu_a) in FP64 */
    float l2s_hi, l2s_lo; /* -log2(0.1f): one roulette survival */
    float spm; /* shells per mean free path, FP64 rounded once */
    float shell_max;
    uint32_t last_shell;
    uint32_t key0, key1;
};

__device__ __forceinline__ void launch_photon(PhotonW& p, uint64_t photon, uint32_t key0, uint32_t key1)
{
    p.x = p.y = p.z = 0.0f;
    p.xl = p.yl = p.zl = 0.0f;
    p.u = p.v = 0.0f;
    p.w = 1.0f;
    p.weight = 1.0f;
    p.e = 0;
    p.r = 0;
    p.k = 0;
    p.plo = (uint32_t)photon;
    p.phi = (uint32_t)(photon >> 32);
    p.pp = tmc_philox_photon_init(p.plo, p.phi, key0, key1);
}

/* 2^(e * (ahi + alo) + r * (shi + slo)) to within about one float rounding
 * of the exact value (e, r < 2^24 are exact in float). */
__device__ __forceinline__ float weight_from_counts(uint32_t e, uint32_t r, const PhysicsW& ph)
{
    const float fe = (float)e, fr = (float)r;
    const float p1 = fe * ph.l2a_hi;
    const float q1 = __builtin_fmaf(fe, ph.l2a_hi, -p1); /* exact */
    const float p2 = fr * ph.l2s_hi;
    const float q2 = __builtin_fmaf(fr, ph.l2s_hi, -p2); /* exact */
    const float s = p1 + p2;
    const float bb = s - p1;
    const float es = (p1 - (s - bb)) + (p2 - bb); /* two-sum error */
    const float lo = __builtin_fmaf(fe, ph.l2a_lo, fr * ph.l2s_lo);
    const float d = es + (q1 + q2 + lo);
    /* 2^(s + d) = 2^s (1 + d ln 2 + ...), |d| well below 2^-20 */
    return __builtin_amdgcn_exp2f(s) * __builtin_fmaf(d, 0x1.62e430p-1f, 1.0f);
}

/* x += t * u carried as the pair (x, xl). */
__device__ __forceinline__ void compensated_move(float& x, float& xl, float t, float u)
{
    const float d = t * u;
    const float ed = __builtin_fmaf(t, u, -d); /* exact */
    const float s = x + d;
    const float bb = s - x;
    const float es = (x - (s - bb)) + (d - bb); /* exact */
    const float lo = xl + (ed + es);
    x = s + lo;
    xl = lo - (x - s);
}

template <uint32_t SAMPLER, bool COMP>
__device__ __forceinline__ bool step(PhotonW& p, const PhysicsW& ph, const LaneTally& lt)
{
    const tmc_u32x4 r = tmc_philox4x32_10_block0(p.k, p.pp, ph.key0, ph.key1); /* block (k, 0) */

    /* move (photon.c:21-24) */
    const float t = -log32_u24(tmc_u24(r.x0)); /* = -log32(tmc_u01(r.x0)) */
    if (COMP) {
        compensated_move(p.x, p.xl, t, p.u);
        compensated_move(p.y, p.yl, t, p.v);
        compensated_move(p.z, p.zl, t, p.w);
    } else {
        p.x = __builtin_fmaf(t, p.u, p.x);
        p.y = __builtin_fmaf(t, p.v, p.y);
        p.z = __builtin_fmaf(t, p.w, p.z);
    }

    /* absorb (photon.c:26-32) */
    const float r2 = __builtin_fmaf(p.z, p.z, __builtin_fmaf(p.y, p.y, p.x * p.x));
    const float f = sqrt_radius(r2) * ph.spm;
    tally_add(lt, tally_pair(lt, f < ph.shell_max, (uint32_t)f), fx_heat(p.weight), fx_heat2(p.weight));
    ++p.e;
    p.weight = weight_from_counts(p.e, p.r, ph);

    /* new direction (photon.c:34-43), as the FP32 step */
    if (SAMPLER == 1u) {
        uint32_t w0 = r.x1, w1 = r.x2;
        tmc_u32x4 blk = { 0u, 0u, 0u, 0u };
        float xi1, xi2, tt;
        uint32_t att = 0;
        for (;;) {
            xi1 = __builtin_fmaf(tmc_u24(w0), 0x1p-23f, -1.0f);
            xi2 = __builtin_fmaf(tmc_u24(w1), 0x1p-23f, -1.0f);
            tt = __builtin_fmaf(xi2, xi2, xi1 * xi1);
            ++att;
            if ((tt <= 1.0f && tt > 0.0f) || att >= 64u)
                break;
            if (att & 1u) {
                blk = tmc_philox4x32_10(p.k, (att + 1u) >> 1, p.plo, p.phi, ph.key0, ph.key1);
                w0 = blk.x0;
                w1 = blk.x1;
            } else {
                w0 = blk.x2;
                w1 = blk.x3;
            }
        }
        p.u = __builtin_fmaf(2.0f, tt, -1.0f);
        const float s = 2.0f * sqrt_cr_walk(1.0f - tt); /* 1 - tt: 0 or >= 2^-24 */
        p.v = xi1 * s;
        p.w = xi2 * s;
    } else {
        const float mu = __builtin_fmaf(tmc_u24(r.x1), 0x1p-23f, -1.0f); /* = 2 u - 1, both exact */
        const float s = sqrt_cr_walk(__builtin_fmaf(-mu, mu, 1.0f)); /* 0 or >= 2^-23 */
        float sn, cs;
        sincos2pi32(tmc_u01(r.x2), &sn, &cs);
        p.u = mu;
        p.v = s * cs;
        p.w = s * sn;
    }
    ++p.k;

    /* roulette (photon.c:45-49): a survival multiplies by 1/0.1f */
    if (p.weight < 0.001f) {
        if (tmc_u01(r.x3) > 0.1f)
            return false;
        ++p.r;
        p.weight = weight_from_counts(p.e, p.r, ph);
    }
    return true;
}

/* ---------------------------------------------------------------------- */
/* FP64                                                                    */
/* ---------------------------------------------------------------------- */

struct Photon64 {
    double x, y, z, u, v, w, weight;
    uint32_t k;
    uint32_t plo, phi;
    tmc_philox_photon pp;
};

struct Physics64 {
    double albedo; /* mu_s / (mu_s + mu_a) in FP64 */
    double spm; /* 1e4 / microns_per_shell / (mu_a + mu_s) in FP64 */
    double shell_max;
    uint32_t last_shell;
    uint32_t key0, key1;
};

__device__ __forceinline__ void launch_photon(Photon64& p, uint64_t photon, uint32_t key0, uint32_t key1)
{
    p.x = p.y = p.z = 0.0;
    p.u = p.v = 0.0;
    p.w = 1.0;
    p.weight = 1.0;
    p.k = 0;
    p.plo = (uint32_t)photon;
    p.phi = (uint32_t)(photon >> 32);
    p.pp = tmc_philox_photon_init(p.plo, p.phi, key0, key1);
}

__device__ __forceinline__ uint64_t fx_heat64(double w) { return (uint64_t)__double2ull_rn(w * 0x1p33); }
__device__ __forceinline__ uint64_t fx_heat2_64(double w) { return (uint64_t)__double2ull_rn((w * w) * 0x1p43); }

template <uint32_t SAMPLER>
__device__ __forceinline__ bool step(Photon64& p, const Physics64& ph, const LaneTally& lt)
{
    const tmc_u32x4 r = tmc_philox4x32_10_block0(p.k, p.pp, ph.key0, ph.key1); /* block (k, 0) */

    /* move (photon.c:21-24) */
    const double t = -log((double)tmc_u01(r.x0));
    p.x += t * p.u;
    p.y += t * p.v;
    p.z += t * p.w;

    /* absorb (photon.c:26-32) */
    const double f = sqrt(p.x * p.x + p.y * p.y + p.z * p.z) * ph.spm;
    tally_add(lt, tally_pair(lt, f < ph.shell_max, (uint32_t)f), fx_heat64(p.weight), fx_heat2_64(p.weight));
    p.weight *= ph.albedo;

    /* new direction (photon.c:34-43) */
    if (SAMPLER == 1u) {
        uint32_t w0 = r.x1, w1 = r.x2;
        tmc_u32x4 blk = { 0u, 0u, 0u, 0u };
        double xi1, xi2, tt;
        uint32_t att = 0;
        for (;;) {
            xi1 = 2.0 * (double)tmc_u01(w0) - 1.0;
            xi2 = 2.0 * (double)tmc_u01(w1) - 1.0;
            tt = xi1 * xi1 + xi2 * xi2;
            ++att;
            if ((tt <= 1.0 && tt != 0.0) || att >= 64u)
                break;
            if (att & 1u) {
                blk = tmc_philox4x32_10(p.k, (att + 1u) >> 1, p.plo, p.phi, ph.key0, ph.key1);
                w0 = blk.x0;
                w1 = blk.x1;
            } else {
                w0 = blk.x2;
                w1 = blk.x3;
            }
        }
        p.u = 2.0 * tt - 1.0;
        p.v = xi1 * sqrt((1.0 - p.u * p.u) / tt);
        p.w = xi2 * sqrt((1.0 - p.u * p.u) / tt);
    } else {
        const double mu = 2.0 * (double)tmc_u01(r.x1) - 1.0;
        const double s = sqrt(1.0 - mu * mu);
        double sn, cs;
        sincos(2.0 * M_PI * (double)tmc_u01(r.x2), &sn, &cs);
        p.u = mu;
        p.v = s * cs;
        p.w = s * sn;
    }
    ++p.k;

    /* roulette (photon.c:45-49), the float literals widened as in C */
    if (p.weight < (double)0.001f) {
        if ((double)tmc_u01(r.x3) > (double)0.1f)
            return false;
        p.weight /= (double)0.1f;
    }
    return true;
}

} // namespace tmc

#endif
