/* Every entry point of the oracle (oracle/oracle.c) on small and edge-case
 * inputs, built with AddressSanitizer and UndefinedBehaviorSanitizer: one
 * shell, shells thinner than a step, a low albedo whose deposits fall below
 * the fixed-point range, a high albedo, both samplers, interaction caps,
 * several threads, more threads than photons, zero photons, and bounded
 * text buffers down to none.  Besides the sanitizers' own checks it asserts
 * a few invariants that hold for any input (energy balance of the FP64
 * walk, the mirror's tally against the FP64 heat, the table's length).
 * Test infrastructure; prints "oracle_check: ok" at the end. */
#include "oracle.c"

#define CHECK(c)                                                                   \
    do {                                                                           \
        if (!(c)) {                                                                \
            fprintf(stderr, "oracle_check: %s:%d: %s\n", __FILE__, __LINE__, #c); \
            exit(1);                                                               \
        }                                                                          \
    } while (0)

static const orc_params cases[] = {
    { 2.0f, 20.0f, 101, 50.0f }, /* params.h defaults */
    { 2.0f, 20.0f, 1, 50.0f }, /* one shell: everything in the last one */
    { 2.0f, 20.0f, 7, 0.5f }, /* shells much thinner than a step */
    { 19.0f, 1.0f, 13, 100.0f }, /* albedo 0.05: truncated fixed-point images */
    { 10.0f, 10.0f, 51, 100.0f }, /* albedo 0.5 */
    { 0.5f, 50.0f, 201, 20.0f }, /* albedo 0.99 */
    { 1.0f, 0.0f, 3, 1e6f }, /* no scattering, one huge shell */
};

static void check_walks(const orc_params* p)
{
    const size_t S = p->shells;
    double* h = calloc(S, sizeof(double));
    double* h2 = calloc(S, sizeof(double));
    uint64_t* tally = calloc(4 * S, sizeof(uint64_t));
    float* fh = calloc(S, sizeof(float));
    float* fh2 = calloc(S, sizeof(float));
    CHECK(h && h2 && tally && fh && fh2);
    float albedo, spm;
    orc_derived32(p, &albedo, &spm);
    CHECK(albedo >= 0.0f && albedo < 1.0f && spm > 0.0f);

    for (int sampler = 0; sampler < 2; ++sampler) {
        for (int threads = 1; threads <= 3; threads += 2) {
            const uint64_t n = 300;
            uint64_t k64 = 0, k32 = 0;
            CHECK(orc_walk64(p, 7, 11, n, sampler, 0, threads, h, h2, &k64) == 0);
            /* every photon deposits its whole unit weight but what the
             * roulette removes or adds, which averages out; the total is
             * within a loose band of n */
            double total = 0.0;
            for (size_t s = 0; s < S; ++s) {
                CHECK(h[s] >= 0.0 && h2[s] >= 0.0 && h2[s] <= h[s] + 1e-12);
                total += h[s];
            }
            CHECK(total > 0.5 * (double)n && total < 1.5 * (double)n);
            CHECK(orc_mirror32(p, 7, 11, n, sampler, 0, threads, tally, &k32) == 0);
            /* the mirror's interactions and deposits follow the FP64 walk
             * closely (the two differ in rounding only) */
            CHECK(k32 > 0 && (double)k32 > 0.9 * (double)k64 && (double)k32 < 1.1 * (double)k64);
            double total32 = 0.0;
            for (size_t s = 0; s < S; ++s) {
                CHECK(tally[2 * s + 1] == 0); /* 300 photons: no carry into the high word */
                total32 += (1.0 - (double)albedo) * ldexp((double)tally[2 * s], -33);
            }
            CHECK(fabs(total32 - total) < 0.1 * (double)n);
            /* capped walks */
            CHECK(orc_walk64(p, 7, 11, n, sampler, 3, threads, h, h2, &k64) == 0 && k64 <= 3 * n);
            CHECK(orc_mirror32(p, 7, 11, n, sampler, 3, threads, tally, &k32) == 0 && k32 <= 3 * n);
        }
        /* zero photons, and more threads than photons */
        uint64_t k = 1;
        CHECK(orc_walk64(p, 1, 0, 0, sampler, 0, 4, h, h2, &k) == 0 && k == 0);
        CHECK(orc_walk64(p, 1, 0, 2, sampler, 0, 64, h, h2, &k) == 0);
        CHECK(orc_mirror32(p, 1, 0, 2, sampler, 0, 64, tally, &k) == 0);
        /* trace with a short and a long buffer */
        uint32_t shell[64];
        double w[64];
        const uint64_t m = orc_trace64(p, 3, 5, sampler, 64, shell, w);
        CHECK(m >= 1);
        for (uint64_t i = 0; i < (m < 64 ? m : 64); ++i)
            CHECK(shell[i] < p->shells && w[i] >= 0.0 && w[i] <= 1.0);
        (void)orc_trace64(p, 3, 5, sampler, 1, shell, w);
        double uvw[3 * 8];
        CHECK(orc_directions64(9, 2, sampler, 8, uvw) == 0);
        for (int i = 0; i < 8; ++i) {
            const double r = uvw[3 * i] * uvw[3 * i] + uvw[3 * i + 1] * uvw[3 * i + 1] + uvw[3 * i + 2] * uvw[3 * i + 2];
            CHECK(r > 1.0 - 1e-9 && r < 1.0 + 1e-9);
        }
    }
    /* the reference restatement (glibc rand) and its table */
    uint64_t kr = 0;
    CHECK(orc_restate(p, 5, 200, fh, fh2, &kr) == 0 && kr > 0);
    const long len = orc_restate_table(p, 5, 200, NULL, 0);
    CHECK(len > 0);
    char* buf = malloc((size_t)len + 1);
    CHECK(buf);
    CHECK(orc_restate_table(p, 5, 200, buf, (size_t)len + 1) == len && strlen(buf) == (size_t)len);
    char small[17];
    CHECK(orc_restate_table(p, 5, 200, small, sizeof small) == len && strlen(small) == sizeof small - 1);
    CHECK(orc_restate_table(p, 5, 200, small, 1) == len && small[0] == '\0');
    free(buf);
    free(h);
    free(h2);
    free(tally);
    free(fh);
    free(fh2);
}

int main(void)
{
    for (size_t i = 0; i < sizeof cases / sizeof cases[0]; ++i)
        check_walks(&cases[i]);

    /* invalid parameters are refused before any work */
    const orc_params bad[] = { { 0.0f, 1.0f, 3, 1.0f }, { 1.0f, -1.0f, 3, 1.0f }, { 1.0f, 1.0f, 0, 1.0f },
                               { 1.0f, 1.0f, 3, 0.0f }, { NAN, 1.0f, 3, 1.0f }, { 1e-30f, 1.0f, 3, 1.0f } };
    double h[4], h2[4];
    uint64_t t[16], k;
    for (size_t i = 0; i < sizeof bad / sizeof bad[0]; ++i) {
        CHECK(orc_walk64(&bad[i], 1, 0, 4, 0, 0, 1, h, h2, &k) < 0);
        CHECK(orc_mirror32(&bad[i], 1, 0, 4, 0, 0, 1, t, &k) < 0);
        CHECK(orc_restate_table(&bad[i], 1, 4, NULL, 0) < 0);
    }
    CHECK(orc_walk64(&cases[0], 1, 0, 4, 2, 0, 1, h, h2, &k) < 0); /* unknown sampler */

    /* elementary functions over the uniforms' extremes */
    const float in[] = { 0x1p-24f, 0x1p-23f, 0.25f, 0.5f, 0.75f, 1.0f - 0x1p-24f, 1.0f };
    float lg[7], sn[7], cs[7];
    orc_eval_math32(in, 7, lg, sn, cs);
    for (int i = 0; i < 7; ++i)
        CHECK(lg[i] <= 0.0f && sn[i] >= -1.0f && sn[i] <= 1.0f && cs[i] >= -1.0f && cs[i] <= 1.0f);
    CHECK(orc_log32(1.0f) == 0.0f);
    CHECK(orc_u01_f(0) > 0.0f && orc_u01_f(0xFFFFFFFFu) <= 1.0f);
    const uint32_t ctr[4] = { 0, 0, 0, 0 }, key[2] = { 0, 0 };
    uint32_t out[4];
    orc_philox(ctr, key, out);
    CHECK(out[0] == 0x6627e8d5u && out[1] == 0xe169c58du && out[2] == 0xbc57ac4cu && out[3] == 0x9b00dbd8u);
    puts("oracle_check: ok");
    return 0;
}
