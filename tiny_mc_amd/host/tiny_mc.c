/* tiny_mc driver for the MI355X path: 1 W point source heating an infinite
 * isotropic scattering medium (Prahl's tiny_mc).
 *
 * Drop-in for the reference driver (tiny_mc.c:34-69): same params.h macros,
 * same '#'-prefixed header, timing lines and shell table on stdout
 * (tiny_mc.c:37-40, 55-66).  The srand(SEED) + photon() loop
 * (tiny_mc.c:43-49) becomes one tiny_mc_gpu_run() call, and the tallies
 * arrive in double, so the normalisation is done in double throughout.
 *
 * Build: make -C tiny_mc_amd tiny_mc (optionally with -D overrides in
 * DRIVER_DEFS, e.g. DRIVER_DEFS='-DSEED=12345 -DPHOTONS=1073741824';
 * -DSAMPLER=1 selects the reference's rejection sampler).
 */
#define _XOPEN_SOURCE 500 /* M_PI */

#include "params.h"
#include "wtime.h"

#include "tiny_mc_gpu.h"

#include <assert.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static const char t1[] = "Tiny Monte Carlo by Scott Prahl (http://omlc.ogi.edu)";
static const char t2[] = "1 W Point Source Heating in Infinite Isotropic Scattering Medium";
static const char t3[] = "MI355X version: HIP gfx950 walk kernel behind the tiny_mc_gpu C ABI";

static double heat[SHELLS];
static double heat2[SHELLS];
#if SAMPLER != 0
static uint64_t tally[4 * SHELLS];
#endif

int main(void)
{
    const unsigned long long photons = (unsigned long long)(PHOTONS);

    printf("# %s\n# %s\n# %s\n", t1, t2, t3);
    printf("# Scattering = %8.3f/cm\n", (double)(MU_S));
    printf("# Absorption = %8.3f/cm\n", (double)(MU_A));
    printf("# Photons    = %8llu\n#\n", photons);

    const tmc_params params = { MU_A, MU_S, SHELLS, (float)(MICRONS_PER_SHELL), photons };
    const unsigned long long seed = (unsigned long long)(SEED);

    double start = wtime();
#if SAMPLER != 0
    /* the reference's rejection sampler: the call with options, then the
     * integer tallies in the reference's units */
    tmc_options opt;
    tmc_options_default(&opt);
    opt.sampler = (uint32_t)(SAMPLER);
    int rc = tiny_mc_gpu_run_ex(&params, seed, &opt, tally, NULL);
    if (rc == TMC_OK)
        rc = tiny_mc_gpu_tally_to_heat(&params, tally, heat, heat2);
#else
    int rc = tiny_mc_gpu_run(&params, seed, heat, heat2);
#endif
    double end = wtime();
    if (rc != TMC_OK) {
        fprintf(stderr, "tiny_mc: %s\n", tiny_mc_gpu_strerror(rc));
        return 1;
    }
    assert(start <= end);
    double elapsed = end - start;

    printf("# %lf seconds\n", elapsed);
    printf("# %lf K photons per second\n", 1e-3 * (double)photons / elapsed);

    printf("# Radius\tHeat\n");
    printf("# [microns]\t[W/cm^3]\tError\n");
    const double mps = (double)(MICRONS_PER_SHELL);
    const double t = 4.0 * M_PI * mps * mps * mps * (double)photons / 1e12;
    for (unsigned int i = 0; i < SHELLS - 1; ++i) {
        const double vol = (double)i * i + i + 1.0 / 3.0;
        printf("%6.0f\t%12.5f\t%12.5f\n", i * mps, heat[i] / t / vol,
               sqrt(heat2[i] - heat[i] * heat[i] / (double)photons) / t / vol);
    }
    printf("# extra\t%12.5f\n", heat[SHELLS - 1] / (double)photons);
    return 0;
}
