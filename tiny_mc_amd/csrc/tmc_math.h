/* FP32 elementary functions of the walk, written for gfx950.
 *
 * The walk needs log (step length, photon.c:21), sqrt (radius and direction,
 * photon.c:26,42-43) and, for the analytic direction sampler, sin/cos of a
 * uniform angle.  They are implemented here from fused multiply-adds and
 * exact hardware primitives (frexp, round-to-nearest-even, IEEE sqrt and
 * division), so that the host mirror that uses the same coefficients and
 * operation order reproduces them bit for bit.
 *
 * log:     x = f * 2^e with f in [sqrt(1/2), sqrt(2)), g = f - 1 (exact by
 *          Sterbenz), log x = e ln2 + g + g^2 Q(g); Q is a degree-8 discrete
 *          minimax polynomial (gen_poly.py).  Max error 0.87 ulp over every
 *          input the walk can produce (u = m 2^-24, m = 1..2^24).
 * sincos:  2 pi a with a in [0, 1]: n = rint(4a), r = a - n/4 (exact),
 *          sin(2 pi r) = r S(r^2) (degree-3 S), cos(2 pi r) = C(r^2)
 *          (degree-4 C), then the quadrant rotation.  Max abs error 8.4e-8.
 */
#ifndef TMC_MATH_H
#define TMC_MATH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tiny_mc_gpu.h"
#include "tmc_philox.h"

namespace tmc {

#define TMC_LN2_HI 0x1.62e400p-1f /* 17 significant bits: e * LN2_HI is exact */
#define TMC_LN2_LO 0x1.7f7d1cp-20f
#define TMC_SQRT_HALF 0x1.6a09e6p-1f

/* log32 of x = m * 2^-24 for an integer-valued float m >= 1, taking m:
 * frexp of x and of m differ only in the exponent (by 24), so this equals
 * log32(m * 0x1p-24f) bit for bit without the scaling. */
__device__ __forceinline__ float log32_u24(float m)
{
    int e = __builtin_amdgcn_frexp_expf(m) - 24;
    float f = __builtin_amdgcn_frexp_mantf(m);
    if (f < TMC_SQRT_HALF) {
        f = f + f;
        e -= 1;
    }
    const float g = f - 1.0f;
    float q = -0x1.38301cp-4f;
    q = __builtin_fmaf(q, g, 0x1.084864p-3f);
    q = __builtin_fmaf(q, g, -0x1.0f3766p-3f);
    q = __builtin_fmaf(q, g, 0x1.22719cp-3f);
    q = __builtin_fmaf(q, g, -0x1.54276ep-3f);
    q = __builtin_fmaf(q, g, 0x1.99a3f2p-3f);
    q = __builtin_fmaf(q, g, -0x1.000426p-2f);
    q = __builtin_fmaf(q, g, 0x1.555550p-2f);
    q = __builtin_fmaf(q, g, -0x1.fffff8p-2f);
    const float l = __builtin_fmaf(g * g, q, g);
    const float ef = (float)e;
    return __builtin_fmaf(ef, TMC_LN2_HI, __builtin_fmaf(ef, TMC_LN2_LO, l));
}

__device__ __forceinline__ float log32(float x)
{
    int e = __builtin_amdgcn_frexp_expf(x);
    float f = __builtin_amdgcn_frexp_mantf(x);
    if (f < TMC_SQRT_HALF) {
        f = f + f;
        e -= 1;
    }
    const float g = f - 1.0f;
    float q = -0x1.38301cp-4f;
    q = __builtin_fmaf(q, g, 0x1.084864p-3f);
    q = __builtin_fmaf(q, g, -0x1.0f3766p-3f);
    q = __builtin_fmaf(q, g, 0x1.22719cp-3f);
    q = __builtin_fmaf(q, g, -0x1.54276ep-3f);
    q = __builtin_fmaf(q, g, 0x1.99a3f2p-3f);
    q = __builtin_fmaf(q, g, -0x1.000426p-2f);
    q = __builtin_fmaf(q, g, 0x1.555550p-2f);
    q = __builtin_fmaf(q, g, -0x1.fffff8p-2f);
    const float l = __builtin_fmaf(g * g, q, g);
    const float ef = (float)e;
    return __builtin_fmaf(ef, TMC_LN2_HI, __builtin_fmaf(ef, TMC_LN2_LO, l));
}

__device__ __forceinline__ void sincos2pi32(float a, float* s_out, float* c_out)
{
    const float n = __builtin_rintf(a * 4.0f);
    const float r = __builtin_fmaf(n, -0.25f, a);
    const float r2 = r * r;
    float ps = -0x1.2d9302p+6f;
    ps = __builtin_fmaf(ps, r2, 0x1.465e92p+6f);
    ps = __builtin_fmaf(ps, r2, -0x1.4abbbap+5f);
    ps = __builtin_fmaf(ps, r2, 0x1.921fb6p+2f);
    const float sn = r * ps;
    float pc = 0x1.d9d57ep+5f;
    pc = __builtin_fmaf(pc, r2, -0x1.55c5e0p+6f);
    pc = __builtin_fmaf(pc, r2, 0x1.03c1dep+6f);
    pc = __builtin_fmaf(pc, r2, -0x1.3bd3ccp+4f);
    pc = __builtin_fmaf(pc, r2, 1.0f);
    const int q = (int)n;
    const bool swap = (q & 1) != 0;
    const float s0 = swap ? pc : sn;
    const float c0 = swap ? sn : pc;
    /* quadrant 0: ( s,  c)  1: ( c, -s)  2: (-s, -c)  3: (-c,  s) */
    *s_out = (q & 2) ? -s0 : s0;
    *c_out = ((q + 1) & 2) ? -c0 : c0;
}

/* The two transcendental quantities of a step, by math mode:
 *   TMC_MATH_HARDWARE: the hardware base-2 log (v_log_f32) scaled by -ln 2,
 *     and the hardware sine and cosine of a turn fraction (v_sin_f32 /
 *     v_cos_f32 take their argument in revolutions);
 *   TMC_MATH_OWN: the polynomials above, which a host program reproduces
 *     without the device.
 * r is the 32-bit Philox word of the uniform u = ((r >> 8) + 1) 2^-24. */
template <uint32_t MATH>
__device__ __forceinline__ float step_length(uint32_t r)
{
    if (MATH == TMC_MATH_HARDWARE)
        return __builtin_amdgcn_logf(tmc_u01_fma(r)) * -0x1.62e430p-1f;
    return -log32_u24((float)((r >> 8) + 1u)); /* = -log32(u) */
}

template <uint32_t MATH>
__device__ __forceinline__ void sincos_turn(uint32_t r, float* s_out, float* c_out)
{
    const float a = tmc_u01_fma(r);
    if (MATH == TMC_MATH_HARDWARE) {
        *s_out = __builtin_amdgcn_sinf(a);
        *c_out = __builtin_amdgcn_cosf(a);
    } else {
        sincos2pi32(a, s_out, c_out);
    }
}

/* Forward declaration: defined below. */
__device__ __forceinline__ float sqrt_cr_walk(float x);

/* sin(theta) = sqrt(1 - mu^2) of the analytic direction, mu = 2u - 1 on the
 * uniform's 2^24-point grid (the radicand, rounded once by the FMA, is 0 or
 * >= 2^-23).  TMC_MATH_HARDWARE takes the hardware square root (v_sqrt_f32,
 * within one ulp): like -ln u and sin/cos it is then a function of the
 * uniform's 24 bits, which a host program reproduces from a table of the
 * device's values.  TMC_MATH_OWN rounds correctly. */
template <uint32_t MATH>
__device__ __forceinline__ float sin_polar(float mu)
{
    const float x = __builtin_fmaf(-mu, mu, 1.0f);
    if (MATH == TMC_MATH_HARDWARE)
        return __builtin_amdgcn_sqrtf(x);
    return sqrt_cr_walk(x);
}

/* Correctly rounded square root (IEEE), the compiler's default lowering of
 * sqrtf under -fhip-fp32-correctly-rounded-divide-sqrt. */
__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }

/* Correctly rounded square root for x == 0 or 2^-96 <= x < inf, which covers
 * every radicand of the walk but the radius of a near-exact cancellation.
 * The hardware estimate (v_sqrt_f32, within one ulp) is corrected by one
 * ulp from the signs of the residuals x - s'*s of its two neighbours s'
 * (exact by FMA): the step of the compiler's IEEE lowering without its
 * scaling of tiny inputs and its infinity / NaN screening (x = 0: the lower
 * neighbour's residual is NaN and the upper one's is 0, so 0 stays).  Equal
 * to sqrt_rn on every float of the domain (exhaustive device check,
 * tiny_mc_gpu_check_sqrt). */
__device__ __forceinline__ float sqrt_cr_walk(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    const float rd = __builtin_fmaf(-sd, s, x);
    const float ru = __builtin_fmaf(-su, s, x);
    const float y = rd <= 0.0f ? sd : s;
    return ru > 0.0f ? su : y;
}

/* Smallest positive radicand sqrt_cr_walk is exact for. */
#define TMC_SQRT_WALK_MIN 0x1p-96f

/* Correctly rounded radius sqrt(r2) of a walk position: the short form
 * unless some lane of the wave has a radicand below 2^-96 (a zero-length
 * first step, probability 2^-24 per photon, or a cancellation to within
 * 2^-48 of the origin, never seen), where the IEEE lowering is used. */
__device__ __forceinline__ float sqrt_radius(float r2)
{
    if (__builtin_expect(__ballot(r2 < TMC_SQRT_WALK_MIN) != 0, 0))
        return sqrt_rn(r2);
    return sqrt_cr_walk(r2);
}

/* w / 0.1f correctly rounded, as the product of w and the real constant
 * C = 1 / 0.1f = 9.99999985...: fma(w, C_hi, w * C_lo) with C_hi = RN(C) = 10
 * and C_lo = RN(C - C_hi) = -0x1.4p-23 (Brisebarre and Muller, "Correctly
 * rounded multiplication by arbitrary precision constants", IEEE TC 2008).
 * For this constant the form equals the IEEE quotient on every float with
 * |w| >= 2^-100 (exhaustive device check, tiny_mc_gpu_check_short_op; the
 * first failures lie below 2^-106, where w * C_lo loses bits to
 * underflow).  One multiply and one FMA instead of the division sequence. */
#define TMC_DIV_TENTH_MIN 0x1p-100f

__device__ __forceinline__ float div_tenth_short(float w) { return __builtin_fmaf(w, 10.0f, w * -0x1.4p-23f); }

/* w / 0.1f for any w: the short form unless some lane of the wave holds a
 * weight of magnitude below 2^-100 (reachable only with albedos below
 * ~1e-27). */
__device__ __forceinline__ float div_tenth(float w)
{
    if (__builtin_expect(__ballot(__builtin_fabsf(w) < TMC_DIV_TENTH_MIN) != 0, 0))
        return w / 0.1f;
    return div_tenth_short(w);
}

/* Fixed-point images of one deposit: w * 2^33 and (w*w) * 2^43, truncated.
 * Exact for every w in [2^-10, 1].  At an albedo of 0.1 or more every
 * deposit lies there (a weight below the 0.001 cutoff is rouletted before
 * the next deposit, photon.c:45-49, and a survivor re-enters at about 10 a
 * times the weight it fell from), so the sums are the exact sums of the
 * FP32 deposits.  Below 0.1 each survival shrinks the weight by about 10 a,
 * and after enough survivals (two at albedo 0.05, about twenty at 0.09)
 * deposits fall below 2^-10 and their images are truncated:
 * deterministically, and the same in the host mirror. */
#define TMC_HEAT_FRAC_BITS 33
#define TMC_HEAT2_FRAC_BITS 43

/* floor(y * 2^32) as a 64-bit integer for 0 <= y < 2^24: the high word is
 * trunc(y) and the low word trunc(fract(y) * 2^32); fract(y) = y - floor(y)
 * (v_fract_f32) and its scaling by 2^32 are exact, so the pair is the
 * truncation of the real y * 2^32.  Three instructions where the generic
 * float -> u64 conversion takes many more. */
__device__ __forceinline__ uint64_t fixed_u64(float y)
{
    const uint32_t hi = (uint32_t)y;
    const uint32_t lo = (uint32_t)(__builtin_amdgcn_fractf(y) * 0x1p32f);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}

/* w * 2^33 = (2 w) * 2^32 and (w*w) * 2^43 = ((w*w) * 2^11) * 2^32; the
 * power-of-two scalings are exact. */
__device__ __forceinline__ uint64_t fx_heat(float w) { return fixed_u64(w * 2.0f); }
__device__ __forceinline__ uint64_t fx_heat2(float w) { return fixed_u64((w * w) * 0x1p11f); }

} // namespace tmc

#endif
