/* tiny_mc_gpu_run and tiny_mc_gpu_strerror over the FP64 oracle, for the
 * sanitised CPU build of the C driver: the driver (tiny_mc_amd/host/tiny_mc.c)
 * is linked against this file instead of libtiny_mc_gpu.so, so its table and
 * error paths run under AddressSanitizer and UndefinedBehaviorSanitizer with
 * no device.  Test infrastructure.
 *
 * TMC_FAKE_STATUS=<n> (n != 0) makes tiny_mc_gpu_run return n. */
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
