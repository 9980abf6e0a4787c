/* The entry points the C driver calls (tiny_mc_gpu_run, or with SAMPLER != 0
 * tmc_options_default + tiny_mc_gpu_run_ex + tiny_mc_gpu_tally_to_heat, and
 * tiny_mc_gpu_strerror) over the oracle, for the sanitised CPU build of the
 * driver: the driver (tiny_mc_amd/host/tiny_mc.c) is linked against this
 * file instead of libtiny_mc_gpu.so, so its table and error paths run under
 * AddressSanitizer and UndefinedBehaviorSanitizer with no device.
 * tiny_mc_gpu_run walks the FP64 oracle; tiny_mc_gpu_run_ex the FP32 mirror
 * (own math) with the requested sampler.  Test infrastructure.
 *
 * TMC_FAKE_STATUS=<n> (n != 0) makes the walks return n. */
#include "oracle.c"

#include "tiny_mc_gpu.h"

int tiny_mc_gpu_run(const tmc_params* p, uint64_t seed, double* heat, double* heat2)
{
    const char* f = getenv("TMC_FAKE_STATUS");
    if (f && atoi(f))
        return atoi(f);
    if (!p || !heat || !heat2)
        return TMC_E_INVALID;
    const orc_params q = { p->mu_a, p->mu_s, p->shells, p->microns_per_shell };
    return orc_walk64(&q, seed, 0, p->photons, ORC_SAMPLER_ANALYTIC, 0, 2, heat, heat2, NULL) == 0 ? TMC_OK
                                                                                                   : TMC_E_INVALID;
}

const char* tiny_mc_gpu_strerror(int status)
{
    return status == TMC_OK ? "ok" : status == TMC_E_NODEVICE ? "no device" : "error";
}

void tmc_options_default(tmc_options* opt)
{
    memset(opt, 0, sizeof *opt);
}

int tiny_mc_gpu_run_ex(const tmc_params* p, uint64_t seed, const tmc_options* opt, uint64_t* tally,
                       tmc_run_stats* stats)
{
    const char* f = getenv("TMC_FAKE_STATUS");
    if (f && atoi(f))
        return atoi(f);
    if (!p || !opt || !tally)
        return TMC_E_INVALID;
    const orc_params q = { p->mu_a, p->mu_s, p->shells, p->microns_per_shell };
    (void)stats;
    return orc_mirror32(&q, seed, opt->first_photon, p->photons, (int)opt->sampler, 0, 2, tally, NULL) == 0
        ? TMC_OK
        : TMC_E_INVALID;
}

int tiny_mc_gpu_tally_to_heat(const tmc_params* p, const uint64_t* tally, double* heat, double* heat2)
{
    if (!p || !tally || !heat || !heat2)
        return TMC_E_INVALID;
    /* the library's conversion: (1 - albedo) in double from the parameters,
     * 128-bit fixed-point sums with TMC_HEAT_FRAC_BITS / TMC_HEAT2_FRAC_BITS */
    const double c = 1.0 - (double)p->mu_s / ((double)p->mu_s + (double)p->mu_a);
    const uint32_t S = p->shells;
    for (uint32_t s = 0; s < S; ++s) {
        const double h = ldexp((double)tally[2 * s + 1], 64) + (double)tally[2 * s];
        const double h2 = ldexp((double)tally[2 * S + 2 * s + 1], 64) + (double)tally[2 * S + 2 * s];
        heat[s] = ldexp(c, -TMC_HEAT_FRAC_BITS) * h;
        heat2[s] = ldexp(c * c, -TMC_HEAT2_FRAC_BITS) * h2;
    }
    return TMC_OK;
}
