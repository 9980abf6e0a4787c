/* tiny_mc oracle — TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the reference's photon transport path, used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in tiny_mc_amd/ links, loads or calls this file.
 *
 * What is here (reference = computacionparalela/tiny_mc):
 *
 *  1. orc_restate / orc_restate_table
 *     The reference loop exactly: glibc srand/rand (tiny_mc.c:43,
 *     photon.c:21,37,38,46), FP32 physics in the reference's operation order
 *     (photon.c:6-51), FP32 accumulation into caller arrays (tiny_mc.c:26-27,
 *     47-49) and the table normalisation and format of tiny_mc.c:37-66.
 *     Pinned byte-for-byte against tables printed by the unmodified reference
 *     (tests/golden/ref_*.txt).
 *
 *  2. orc_walk64
 *     The same physics text (orc_walk_body.h) instantiated in double with
 *     glibc rand() swapped for the Philox4x32-10 stream keyed by (seed,
 *     photon index) that the GPU path uses.  This is the FP64 oracle that GPU
 *     results are compared with.  Pinned by (1) through the shared walk text,
 *     by the Philox known-answer vectors, and by closed forms of the walk
 *     (weight cadence, mean interactions, energy balance, first-step shell
 *     probability, direction moments) in tests/test_oracle_*.py.
 *
 *  3. orc_mirror32 / orc_mirror32_tab
 *     The GPU kernel's exact FP32 arithmetic (same Philox words, same own log
 *     and sin/cos polynomials, fused multiply-adds where the kernel fuses,
 *     correctly rounded sqrt and division) with the kernel's 64-bit
 *     fixed-point tallies, so GPU tallies can be compared bit for bit.  Pinned
 *     against orc_walk64 per photon and per shell, and its log/sin/cos against
 *     libm over every possible input.  For the kernel's hardware-math mode
 *     (v_log_f32, v_sin_f32, v_cos_f32, whose exact rounding is not
 *     published), orc_mirror32_tab takes the step length and sin/cos for each
 *     of the 2^24 uniforms the walk can draw as tables (evaluated by the
 *     device; their accuracy is pinned against libm separately) and restates
 *     every other operation of the walk as above.
 *
 * Build: oracle/Makefile (gcc -std=c11 -ffp-contract=off; no FP contraction,
 * matching the reference's ISO-C build, Makefile:5).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc_philox.h"

#define ORC_SAMPLER_ANALYTIC 0
#define ORC_SAMPLER_REJECTION 1
#define ORC_MAX_ATTEMPTS 64u
#define ORC_MAX_THREADS 256

#define ORC_EXPORT __attribute__((visibility("default")))

typedef struct orc_params {
    float mu_a;              /* absorption, 1/cm (params.h MU_A) */
    float mu_s;              /* scattering, 1/cm (params.h MU_S) */
    uint32_t shells;         /* params.h SHELLS; last shell collects the rest */
    float microns_per_shell; /* params.h MICRONS_PER_SHELL */
} orc_params;

static int orc_params_ok(const orc_params* p)
{
    if (!p || !(p->mu_a > 0.0f) || !(p->mu_s >= 0.0f) || !isfinite(p->mu_a) || !isfinite(p->mu_s))
        return 0;
    if (p->shells < 1 || !(p->microns_per_shell > 0.0f) || !isfinite(p->microns_per_shell))
        return 0;
    const float albedo = p->mu_s / (p->mu_s + p->mu_a);
    return albedo < 1.0f;
}

/* ------------------------------------------------------------------------ */
/* 1. Restatement with glibc rand()                                          */
/* ------------------------------------------------------------------------ */

typedef struct glibc_src {
    int unused;
} glibc_src;

typedef struct f32_sink {
    float* heat;
    float* heat2;
    float one_minus_albedo;
} f32_sink;

static inline float glibc_u(void)
{
    return (float)rand() / (float)RAND_MAX; /* photon.c:21,46 (and :37-38 via 2*U-1) */
}

#define W_FN restate_photon
#define W_REAL float
#define W_SRC glibc_src
#define W_SINK f32_sink
#define W_U_STEP(s, k) glibc_u()
#define W_U_DIR(s, k, a, i) glibc_u()
#define W_U_ROUL(s, k) glibc_u()
#define W_LOG logf
#define W_SQRT sqrtf
#define W_SINCOS2PI(a, s, c) (s = 0.0f, c = 0.0f) /* unused: restatement is rejection only */
#define W_DEPOSIT(sink, shell, wt)                                                          \
    do {                                                                                  \
        (sink)->heat[shell] += (sink)->one_minus_albedo * (wt);                           \
        (sink)->heat2[shell] += (sink)->one_minus_albedo * (sink)->one_minus_albedo * (wt) \
            * (wt);                                                                       \
    } while (0)
#include "orc_walk_body.h"
#undef W_FN
#undef W_REAL
#undef W_SRC
#undef W_SINK
#undef W_U_STEP
#undef W_U_DIR
#undef W_U_ROUL
#undef W_LOG
#undef W_SQRT
#undef W_SINCOS2PI
#undef W_DEPOSIT

/* Derived constants exactly as photon.c:8-9 evaluates them: albedo in float,
 * shells_per_mfp through a double intermediate (1e4 is a double literal). */
static float restate_albedo(const orc_params* p) { return p->mu_s / (p->mu_s + p->mu_a); }
static float restate_spm(const orc_params* p)
{
    return (float)(1e4 / (double)p->microns_per_shell / (double)(p->mu_a + p->mu_s));
}

/* The reference loop (tiny_mc.c:43-49): srand(seed), then photon() into
 * float arrays that are accumulated, never zeroed (caller zeroes them). */
ORC_EXPORT int orc_restate(const orc_params* p, unsigned seed, uint32_t photons, float* heat, float* heat2,
                           uint64_t* interactions)
{
    if (!orc_params_ok(p) || !heat || !heat2)
        return -1;
    f32_sink sink = { heat, heat2, 1.0f - restate_albedo(p) };
    glibc_src src = { 0 };
    uint64_t n = 0;
    srand(seed);
    for (uint32_t i = 0; i < photons; ++i)
        n += restate_photon(&src, &sink, restate_albedo(p), restate_spm(p), p->shells, ORC_SAMPLER_REJECTION, 0);
    if (interactions)
        *interactions = n;
    return 0;
}

/* Appends formatted text to a bounded buffer. */
typedef struct sbuf {
    char* buf;
    size_t cap, len;
} sbuf;

static void sb_printf(sbuf* b, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
#include <stdarg.h>
static void sb_printf(sbuf* b, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    const size_t room = b->len < b->cap ? b->cap - b->len : 0;
    const int n = vsnprintf(b->buf && room ? b->buf + b->len : NULL, room, fmt, ap);
    va_end(ap);
    if (n > 0)
        b->len += (size_t)n;
}

/* The reference's stdout (tiny_mc.c:37-66) without the two timing lines
 * (tiny_mc.c:55-56), with the table arithmetic in the reference's types:
 * t is a float computed through doubles (tiny_mc.c:60); the heat column
 * divides float by float, then by a double (tiny_mc.c:63); the error column
 * subtracts in float and normalises by a float shell factor (tiny_mc.c:64).
 * Returns the text length (excluding the terminator) or a negative error. */
ORC_EXPORT long orc_restate_table(const orc_params* p, unsigned seed, uint32_t photons, char* buf, size_t cap)
{
    if (!orc_params_ok(p) || photons > 0x7fffffffu)
        return -1;
    float* heat = calloc(p->shells, sizeof(float));
    float* heat2 = calloc(p->shells, sizeof(float));
    if (!heat || !heat2) {
        free(heat);
        free(heat2);
        return -2;
    }
    orc_restate(p, seed, photons, heat, heat2, NULL);

    sbuf b = { buf, cap, 0 };
    const int np = (int)photons;
    sb_printf(&b, "# %s\n# %s\n# %s\n", "Tiny Monte Carlo by Scott Prahl (http://omlc.ogi.edu)",
              "1 W Point Source Heating in Infinite Isotropic Scattering Medium",
              "CPU version, adapted for PEAGPGPU by Gustavo Castellano and Nicolas Wolovick");
    sb_printf(&b, "# Scattering = %8.3f/cm\n", p->mu_s);
    sb_printf(&b, "# Absorption = %8.3f/cm\n", p->mu_a);
    sb_printf(&b, "# Photons    = %8d\n#\n", np);
    sb_printf(&b, "# Radius\tHeat\n");
    sb_printf(&b, "# [microns]\t[W/cm^3]\tError\n");
    const float mps = p->microns_per_shell;
    const float t = 4.0f * M_PI * powf(mps, 3.0f) * np / 1e12;
    for (unsigned int i = 0; i < p->shells - 1; ++i) {
        sb_printf(&b, "%6.0f\t%12.5f\t%12.5f\n", i * mps, heat[i] / t / (i * i + i + 1.0 / 3.0),
                  sqrt(heat2[i] - heat[i] * heat[i] / np) / t / (i * i + i + 1.0f / 3.0f));
    }
    sb_printf(&b, "# extra\t%12.5f\n", heat[p->shells - 1] / np);
    free(heat);
    free(heat2);
    if (buf && cap)
        buf[b.len < cap ? b.len : cap - 1] = '\0';
    return (long)b.len;
}

/* ------------------------------------------------------------------------ */
/* 2. FP64 oracle on the Philox stream                                       */
/* ------------------------------------------------------------------------ */

typedef struct phx_src {
    uint64_t seed, photon;
    uint32_t k, b;
    int valid;
    uint32_t blk[4];
} phx_src;

/* Word i of block b of interaction k (counter (k, b, photon_lo, photon_hi)).
 * Interaction indices are 32-bit in the counter, as on the GPU. */
static inline uint32_t phx_word(phx_src* s, uint64_t k, uint32_t b, uint32_t i)
{
    const uint32_t k32 = (uint32_t)k;
    if (!s->valid || s->k != k32 || s->b != b) {
        orc_block(s->seed, s->photon, k32, b, s->blk);
        s->k = k32;
        s->b = b;
        s->valid = 1;
    }
    return s->blk[i];
}

/* Slot map shared with the GPU kernel:
 *   block (k, 0): word 0 step, words 1-2 direction (attempt 0 / analytic),
 *                 word 3 roulette;
 *   rejection attempt a >= 1: block (k, (a + 1) / 2), words 0-1 for odd a,
 *                 words 2-3 for even a. */
static inline uint32_t phx_dir_word(phx_src* s, uint64_t k, uint32_t a, uint32_t i)
{
    if (a == 0)
        return phx_word(s, k, 0, 1 + i);
    return phx_word(s, k, (a + 1) / 2, ((a & 1u) ? 0u : 2u) + i);
}

typedef struct f64_sink {
    long double* heat;
    long double* heat2;
    double one_minus_albedo;
} f64_sink;

#define W_FN walk64_photon
#define W_REAL double
#define W_SRC phx_src
#define W_SINK f64_sink
#define W_U_STEP(s, k) ((double)orc_u01(phx_word((s), (k), 0, 0)))
#define W_U_DIR(s, k, a, i) ((double)orc_u01(phx_dir_word((s), (k), (a), (i))))
#define W_U_ROUL(s, k) ((double)orc_u01(phx_word((s), (k), 0, 3)))
#define W_LOG log
#define W_SQRT sqrt
#define W_SINCOS2PI(a, s, c) (s = sin(2.0 * M_PI * (a)), c = cos(2.0 * M_PI * (a)))
#define W_DEPOSIT(sink, shell, wt)                                                      \
    do {                                                                              \
        const double d_ = (sink)->one_minus_albedo * (wt);                            \
        (sink)->heat[shell] += d_;                                                    \
        (sink)->heat2[shell] += (sink)->one_minus_albedo * (sink)->one_minus_albedo * \
            (wt) * (wt);                                                              \
    } while (0)
#include "orc_walk_body.h"
#undef W_FN
#undef W_REAL
#undef W_SRC
#undef W_SINK
#undef W_U_STEP
#undef W_U_DIR
#undef W_U_ROUL
#undef W_LOG
#undef W_SQRT
#undef W_SINCOS2PI
#undef W_DEPOSIT

typedef struct job64 {
    const orc_params* p;
    uint64_t seed, first, count, max_k;
    int sampler;
    long double *heat, *heat2;
    uint64_t interactions;
} job64;

static void* run64(void* arg)
{
    job64* j = arg;
    const double albedo = (double)j->p->mu_s / ((double)j->p->mu_s + (double)j->p->mu_a);
    const double spm = 1e4 / (double)j->p->microns_per_shell / ((double)j->p->mu_a + (double)j->p->mu_s);
    f64_sink sink = { j->heat, j->heat2, 1.0 - albedo };
    uint64_t n = 0;
    for (uint64_t i = 0; i < j->count; ++i) {
        phx_src src = { j->seed, j->first + i, 0, 0, 0, { 0, 0, 0, 0 } };
        n += walk64_photon(&src, &sink, albedo, spm, j->p->shells, j->sampler, j->max_k);
    }
    j->interactions = n;
    return NULL;
}

static int clamp_threads(int threads, uint64_t count)
{
    if (threads < 1)
        threads = 1;
    if (threads > ORC_MAX_THREADS)
        threads = ORC_MAX_THREADS;
    if ((uint64_t)threads > count)
        threads = count ? (int)count : 1;
    return threads;
}

/* FP64 oracle over photons [first, first + count) of the stream `seed`.
 * heat[s] = sum of (1 - a) w, heat2[s] = sum of (1 - a)^2 w^2 over deposits
 * in shell s (the reference's units, photon.c:30-31), overwritten. */
ORC_EXPORT int orc_walk64(const orc_params* p, uint64_t seed, uint64_t first, uint64_t count, int sampler,
                          uint64_t max_k, int threads, double* heat, double* heat2, uint64_t* interactions)
{
    if (!orc_params_ok(p) || !heat || !heat2 || (sampler != ORC_SAMPLER_ANALYTIC && sampler != ORC_SAMPLER_REJECTION))
        return -1;
    threads = clamp_threads(threads, count);
    job64 jobs[ORC_MAX_THREADS];
    pthread_t tid[ORC_MAX_THREADS];
    const size_t S = p->shells;
    long double* acc = calloc((size_t)threads * 2 * S, sizeof(long double));
    if (!acc)
        return -2;
    for (int t = 0; t < threads; ++t) {
        const uint64_t lo = count * (uint64_t)t / (uint64_t)threads;
        const uint64_t hi = count * (uint64_t)(t + 1) / (uint64_t)threads;
        jobs[t] = (job64){ p, seed, first + lo, hi - lo, max_k, sampler, acc + (size_t)t * 2 * S,
                           acc + (size_t)t * 2 * S + S, 0 };
    }
    for (int t = 1; t < threads; ++t)
        pthread_create(&tid[t], NULL, run64, &jobs[t]);
    run64(&jobs[0]);
    uint64_t n = jobs[0].interactions;
    for (int t = 1; t < threads; ++t) {
        pthread_join(tid[t], NULL);
        n += jobs[t].interactions;
    }
    for (size_t s = 0; s < S; ++s) {
        long double h = 0, h2 = 0;
        for (int t = 0; t < threads; ++t) {
            h += jobs[t].heat[s];
            h2 += jobs[t].heat2[s];
        }
        heat[s] = (double)h;
        heat2[s] = (double)h2;
    }
    free(acc);
    if (interactions)
        *interactions = n;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3. FP32 mirror of the GPU kernel                                          */
/* ------------------------------------------------------------------------ */

/* Own natural log for x in [2^-24, 1] (the only inputs the walk produces):
 * x = f * 2^e with f in [sqrt(1/2), sqrt(2)), g = f - 1 (exact),
 * log(x) = e*ln2 + g + g^2 * Q(g), Q a degree-8 minimax polynomial
 * (tiny_mc_amd/csrc/gen_poly.py).  Same coefficients and operation order as
 * the device code, so results are bit-identical. */
static const float MQ[9] = { -0x1.fffff8p-2f, 0x1.555550p-2f, -0x1.000426p-2f, 0x1.99a3f2p-3f, -0x1.54276ep-3f,
                             0x1.22719cp-3f,  -0x1.0f3766p-3f, 0x1.084864p-3f, -0x1.38301cp-4f };
#define M_LN2_HI 0x1.62e400p-1f
#define M_LN2_LO 0x1.7f7d1cp-20f
#define M_SQRT_HALF 0x1.6a09e6p-1f

ORC_EXPORT float orc_log32(float x)
{
    int e;
    float f = frexpf(x, &e);
    if (f < M_SQRT_HALF) {
        f = f + f;
        e -= 1;
    }
    const float g = f - 1.0f;
    float q = MQ[8];
    for (int i = 7; i >= 0; --i)
        q = fmaf(q, g, MQ[i]);
    const float l = fmaf(g * g, q, g);
    const float ef = (float)e;
    return fmaf(ef, M_LN2_HI, fmaf(ef, M_LN2_LO, l));
}

/* Own sin and cos of 2*pi*a for a in [0, 1]: quarter-turn reduction
 * n = rint(4a), r = a - n/4 in [-1/8, 1/8] (exact), minimax polynomials
 * sin(2 pi r) = r S(r^2), cos(2 pi r) = C(r^2), then quadrant rotation. */
static const float MS[4] = { 0x1.921fb6p+2f, -0x1.4abbbap+5f, 0x1.465e92p+6f, -0x1.2d9302p+6f };
static const float MC[5] = { 0x1.000000p+0f, -0x1.3bd3ccp+4f, 0x1.03c1dep+6f, -0x1.55c5e0p+6f, 0x1.d9d57ep+5f };

ORC_EXPORT void orc_sincos2pi32(float a, float* s_out, float* c_out)
{
    const float n = rintf(a * 4.0f);
    const float r = fmaf(n, -0.25f, a);
    const float r2 = r * r;
    float ps = MS[3];
    ps = fmaf(ps, r2, MS[2]);
    ps = fmaf(ps, r2, MS[1]);
    ps = fmaf(ps, r2, MS[0]);
    const float sn = r * ps;
    float pc = MC[4];
    pc = fmaf(pc, r2, MC[3]);
    pc = fmaf(pc, r2, MC[2]);
    pc = fmaf(pc, r2, MC[1]);
    pc = fmaf(pc, r2, MC[0]);
    const int q = (int)n & 3;
    float s, c;
    switch (q) {
    case 0: s = sn; c = pc; break;
    case 1: s = pc; c = -sn; break;
    case 2: s = -sn; c = -pc; break;
    default: s = -pc; c = sn; break;
    }
    *s_out = s;
    *c_out = c;
}

ORC_EXPORT float orc_u01_f(uint32_t r) { return orc_u01(r); }

ORC_EXPORT void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    orc_philox4x32_10(ctr, key, out);
}

typedef unsigned __int128 u128;

typedef struct m32_params {
    float albedo, spm;
    uint32_t shells;
    int sampler;
    uint64_t max_k;
    /* NULL: own log / sin / cos and the correctly rounded sin(theta); else
     * per uniform index r >> 8 */
    const float *t_tab, *s_tab, *c_tab, *p_tab;
} m32_params;

/* Fixed-point images of one deposit: w * 2^33 and (w*w) * 2^43, truncated.
 * Exact (no rounding) for every w in [2^-10, 1].  With an albedo of 0.1 or
 * more every deposit lies there (a roulette survivor re-enters at about
 * 10 a times the weight it fell from, 0.001 or more); below 0.1 each
 * survival shrinks the weight by about 10 a, and after enough survivals (two
 * at albedo 0.05, about twenty at 0.09) the images of the deposits are
 * truncated, as the kernel's are. */
static inline uint64_t fx_heat(float w) { return (uint64_t)(w * 0x1p33f); }
static inline uint64_t fx_heat2(float w) { return (uint64_t)((w * w) * 0x1p43f); }

/* One photon in the kernel's FP32 arithmetic. */
static uint64_t mirror_photon(uint64_t seed, uint64_t photon, const m32_params* mp, u128* tw, u128* tw2)
{
    float x = 0.0f, y = 0.0f, z = 0.0f;
    float u = 0.0f, v = 0.0f, w = 1.0f;
    float weight = 1.0f;
    const float shell_max = (float)(mp->shells - 1);
    uint32_t k = 0;
    uint64_t n = 0;
    for (;;) {
        uint32_t b0[4];
        orc_block(seed, photon, k, 0, b0);
        const float t = mp->t_tab ? mp->t_tab[b0[0] >> 8] : -orc_log32(orc_u01(b0[0]));
        x = fmaf(t, u, x);
        y = fmaf(t, v, y);
        z = fmaf(t, w, z);
        const float r2 = fmaf(z, z, fmaf(y, y, x * x));
        const float f = sqrtf(r2) * mp->spm;
        const uint32_t shell = f < shell_max ? (uint32_t)f : mp->shells - 1;
        tw[shell] += fx_heat(weight);
        tw2[shell] += fx_heat2(weight);
        weight = weight * mp->albedo;

        if (mp->sampler == ORC_SAMPLER_REJECTION) {
            float xi1, xi2, tt;
            uint32_t a = 0;
            uint32_t blk[4];
            for (;;) {
                uint32_t r1, r2w;
                if (a == 0) {
                    r1 = b0[1];
                    r2w = b0[2];
                } else {
                    if (a & 1u)
                        orc_block(seed, photon, k, (a + 1) / 2, blk);
                    r1 = blk[(a & 1u) ? 0 : 2];
                    r2w = blk[(a & 1u) ? 1 : 3];
                }
                xi1 = fmaf(2.0f, orc_u01(r1), -1.0f);
                xi2 = fmaf(2.0f, orc_u01(r2w), -1.0f);
                tt = fmaf(xi2, xi2, xi1 * xi1);
                ++a;
                if ((tt <= 1.0f && tt > 0.0f) || a >= ORC_MAX_ATTEMPTS)
                    break;
            }
            u = fmaf(2.0f, tt, -1.0f);
            const float s = 2.0f * sqrtf(1.0f - tt);
            v = xi1 * s;
            w = xi2 * s;
        } else {
            const float mu = fmaf(2.0f, orc_u01(b0[1]), -1.0f);
            const float s = mp->p_tab ? mp->p_tab[b0[1] >> 8] : sqrtf(fmaf(-mu, mu, 1.0f));
            float sn, cs;
            if (mp->s_tab) {
                sn = mp->s_tab[b0[2] >> 8];
                cs = mp->c_tab[b0[2] >> 8];
            } else {
                orc_sincos2pi32(orc_u01(b0[2]), &sn, &cs);
            }
            u = mu;
            v = s * cs;
            w = s * sn;
        }
        ++k;
        ++n;
        if (weight < 0.001f) {
            if (orc_u01(b0[3]) > 0.1f)
                break;
            weight = weight / 0.1f;
        }
        if (mp->max_k && n >= mp->max_k)
            break;
    }
    return n;
}

typedef struct job32 {
    const m32_params* mp;
    uint64_t seed, first, count;
    u128 *tw, *tw2;
    uint64_t interactions;
} job32;

static void* run32(void* arg)
{
    job32* j = arg;
    uint64_t n = 0;
    for (uint64_t i = 0; i < j->count; ++i)
        n += mirror_photon(j->seed, j->first + i, j->mp, j->tw, j->tw2);
    j->interactions = n;
    return NULL;
}

/* Host-side derived constants the GPU path uses (same as photon.c:8-9). */
ORC_EXPORT void orc_derived32(const orc_params* p, float* albedo, float* spm)
{
    *albedo = restate_albedo(p);
    *spm = restate_spm(p);
}

/* FP32 mirror over photons [first, first + count).  tally holds 4*shells
 * 64-bit words: (lo, hi) of the 128-bit sum of w*2^33 for each shell, then
 * (lo, hi) of the 128-bit sum of (w*w)*2^43 for each shell. */
ORC_EXPORT int orc_mirror32_tab(const orc_params* p, uint64_t seed, uint64_t first, uint64_t count, int sampler,
                                uint64_t max_k, int threads, const float* t_tab, const float* s_tab,
                                const float* c_tab, const float* p_tab, uint64_t* tally, uint64_t* interactions)
{
    if (!orc_params_ok(p) || !tally || (sampler != ORC_SAMPLER_ANALYTIC && sampler != ORC_SAMPLER_REJECTION)
        || (t_tab && (!s_tab || !c_tab || !p_tab)))
        return -1;
    m32_params mp = { restate_albedo(p), restate_spm(p), p->shells, sampler, max_k, t_tab, s_tab, c_tab, p_tab };
    threads = clamp_threads(threads, count);
    const size_t S = p->shells;
    u128* acc = calloc((size_t)threads * 2 * S, sizeof(u128));
    if (!acc)
        return -2;
    job32 jobs[ORC_MAX_THREADS];
    pthread_t tid[ORC_MAX_THREADS];
    for (int t = 0; t < threads; ++t) {
        const uint64_t lo = count * (uint64_t)t / (uint64_t)threads;
        const uint64_t hi = count * (uint64_t)(t + 1) / (uint64_t)threads;
        jobs[t] = (job32){ &mp, seed, first + lo, hi - lo, acc + (size_t)t * 2 * S, acc + (size_t)t * 2 * S + S, 0 };
    }
    for (int t = 1; t < threads; ++t)
        pthread_create(&tid[t], NULL, run32, &jobs[t]);
    run32(&jobs[0]);
    uint64_t n = jobs[0].interactions;
    for (int t = 1; t < threads; ++t) {
        pthread_join(tid[t], NULL);
        n += jobs[t].interactions;
    }
    for (size_t s = 0; s < S; ++s) {
        u128 a = 0, b = 0;
        for (int t = 0; t < threads; ++t) {
            a += jobs[t].tw[s];
            b += jobs[t].tw2[s];
        }
        tally[2 * s] = (uint64_t)a;
        tally[2 * s + 1] = (uint64_t)(a >> 64);
        tally[2 * S + 2 * s] = (uint64_t)b;
        tally[2 * S + 2 * s + 1] = (uint64_t)(b >> 64);
    }
    free(acc);
    if (interactions)
        *interactions = n;
    return 0;
}

/* orc_mirror32_tab with the own log / sin / cos polynomials and correctly
 * rounded square roots. */
ORC_EXPORT int orc_mirror32(const orc_params* p, uint64_t seed, uint64_t first, uint64_t count, int sampler,
                            uint64_t max_k, int threads, uint64_t* tally, uint64_t* interactions)
{
    return orc_mirror32_tab(p, seed, first, count, sampler, max_k, threads, NULL, NULL, NULL, NULL, tally,
                            interactions);
}

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                               */
/* ------------------------------------------------------------------------ */

/* Per-photon deposit trace of the FP64 oracle: for the first `cap`
 * interactions of one photon, the shell and the deposited weight w.
 * Returns the number of interactions of the photon (may exceed cap). */
typedef struct trace_sink {
    uint32_t* shell;
    double* w;
    uint64_t n, cap;
} trace_sink;

#define W_FN trace64_photon
#define W_REAL double
#define W_SRC phx_src
#define W_SINK trace_sink
#define W_U_STEP(s, k) ((double)orc_u01(phx_word((s), (k), 0, 0)))
#define W_U_DIR(s, k, a, i) ((double)orc_u01(phx_dir_word((s), (k), (a), (i))))
#define W_U_ROUL(s, k) ((double)orc_u01(phx_word((s), (k), 0, 3)))
#define W_LOG log
#define W_SQRT sqrt
#define W_SINCOS2PI(a, s, c) (s = sin(2.0 * M_PI * (a)), c = cos(2.0 * M_PI * (a)))
#define W_DEPOSIT(sink, sh, wt)                \
    do {                                       \
        if ((sink)->n < (sink)->cap) {         \
            (sink)->shell[(sink)->n] = (sh);   \
            (sink)->w[(sink)->n] = (wt);       \
        }                                      \
        (sink)->n++;                           \
    } while (0)
#include "orc_walk_body.h"
#undef W_FN
#undef W_REAL
#undef W_SRC
#undef W_SINK
#undef W_U_STEP
#undef W_U_DIR
#undef W_U_ROUL
#undef W_LOG
#undef W_SQRT
#undef W_SINCOS2PI
#undef W_DEPOSIT

ORC_EXPORT uint64_t orc_trace64(const orc_params* p, uint64_t seed, uint64_t photon, int sampler, uint64_t cap,
                                uint32_t* shell, double* w)
{
    if (!orc_params_ok(p))
        return 0;
    const double albedo = (double)p->mu_s / ((double)p->mu_s + (double)p->mu_a);
    const double spm = 1e4 / (double)p->microns_per_shell / ((double)p->mu_a + (double)p->mu_s);
    trace_sink sink = { shell, w, 0, cap };
    phx_src src = { seed, photon, 0, 0, 0, { 0, 0, 0, 0 } };
    trace64_photon(&src, &sink, albedo, spm, p->shells, sampler, 0);
    return sink.n;
}

/* Directions drawn by the FP64 oracle's samplers for interactions
 * [0, n) of one photon's stream (3 doubles each), for moment tests. */
ORC_EXPORT int orc_directions64(uint64_t seed, uint64_t photon, int sampler, uint64_t n, double* uvw)
{
    phx_src src = { seed, photon, 0, 0, 0, { 0, 0, 0, 0 } };
    for (uint64_t k = 0; k < n; ++k) {
        double u, v, w;
        if (sampler == ORC_SAMPLER_REJECTION) {
            double xi1, xi2, t;
            uint32_t a = 0;
            do {
                xi1 = 2.0 * (double)orc_u01(phx_dir_word(&src, k, a, 0)) - 1.0;
                xi2 = 2.0 * (double)orc_u01(phx_dir_word(&src, k, a, 1)) - 1.0;
                t = xi1 * xi1 + xi2 * xi2;
                ++a;
            } while ((1.0 < t || t == 0) && a < ORC_MAX_ATTEMPTS);
            u = 2.0 * t - 1.0;
            v = xi1 * sqrt((1.0 - u * u) / t);
            w = xi2 * sqrt((1.0 - u * u) / t);
        } else {
            const double mu = 2.0 * (double)orc_u01(phx_dir_word(&src, k, 0, 0)) - 1.0;
            const double s = sqrt(1.0 - mu * mu);
            const double ph = 2.0 * M_PI * (double)orc_u01(phx_dir_word(&src, k, 0, 1));
            u = mu;
            v = s * cos(ph);
            w = s * sin(ph);
        }
        uvw[3 * k] = u;
        uvw[3 * k + 1] = v;
        uvw[3 * k + 2] = w;
    }
    return 0;
}

/* Exhaustive accuracy of the own log and sin/cos over every uniform the
 * walk can produce (u = m * 2^-24, m = 1 .. 2^24): max error in units of
 * the last place of the exact result, and the max absolute error. */
ORC_EXPORT void orc_math32_errors(double* log_max_ulp, double* sincos_max_abs, double* sincos_max_ulp)
{
    double lmax = 0, smax = 0, sulp = 0;
    for (uint32_t m = 1; m <= (1u << 24); ++m) {
        const float x = (float)m * 0x1p-24f;
        const double ex = log((double)x);
        const double got = orc_log32(x);
        if (ex != 0.0) {
            int e;
            frexp(ex, &e);
            const double ulp = ldexp(1.0, e - 24);
            const double err = fabs(got - ex) / ulp;
            if (err > lmax)
                lmax = err;
        } else if (got != 0.0) {
            lmax = 1e30;
        }
        float s, c;
        orc_sincos2pi32(x, &s, &c);
        const double es = sin(2.0 * M_PI * (double)x), ec = cos(2.0 * M_PI * (double)x);
        const double ds = fabs(s - es), dc = fabs(c - ec);
        if (ds > smax)
            smax = ds;
        if (dc > smax)
            smax = dc;
        int e;
        if (es != 0.0) {
            frexp(es, &e);
            const double u1 = ds / ldexp(1.0, e - 24);
            if (u1 > sulp && fabs(es) > 1e-6)
                sulp = u1;
        }
        if (ec != 0.0) {
            frexp(ec, &e);
            const double u2 = dc / ldexp(1.0, e - 24);
            if (u2 > sulp && fabs(ec) > 1e-6)
                sulp = u2;
        }
    }
    *log_max_ulp = lmax;
    *sincos_max_abs = smax;
    *sincos_max_ulp = sulp;
}

/* The mirror's log / sin(2 pi a) / cos(2 pi a) over an array (for the
 * device-vs-mirror comparison of the elementary functions). */
ORC_EXPORT void orc_eval_math32(const float* in, uint64_t n, float* lg, float* sn, float* cs)
{
    for (uint64_t i = 0; i < n; ++i) {
        lg[i] = orc_log32(in[i]);
        orc_sincos2pi32(in[i], &sn[i], &cs[i]);
    }
}

