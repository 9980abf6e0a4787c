/* Compile-time configuration of the tiny_mc driver.
 *
 * Same macro names, defaults and #ifndef overrides as the reference's
 * params.h:5-27, so existing -D build lines keep working.  The values are
 * copied into a tmc_params at run time, so the GPU path itself takes them as
 * run-time parameters. */
#pragma once

#include <time.h>

#ifndef SHELLS
#define SHELLS 101 /* spherical shells; the last one collects the rest */
#endif

#ifndef PHOTONS
#define PHOTONS 32768 /* photons launched */
#endif

#ifndef MU_A
#define MU_A 2.0f /* absorption coefficient, 1/cm; must be > 0 */
#endif

#ifndef MU_S
#define MU_S 20.0f /* scattering coefficient, 1/cm (isotropic, g = 0) */
#endif

#ifndef MICRONS_PER_SHELL
#define MICRONS_PER_SHELL 50 /* shell thickness, microns */
#endif

#ifndef SEED
#define SEED (time(NULL)) /* stream seed */
#endif

/* Not in the reference: the direction sampler of the GPU walk.  0 (default)
 * draws cos(theta) and phi directly; 1 is the reference's rejection method
 * (photon.c:34-43).  Both give the same isotropic distribution. */
#ifndef SAMPLER
#define SAMPLER 0
#endif
