/* Link-time replacement of glibc srand()/rand() for the UNMODIFIED reference
 * sources (test infrastructure; linked with -Wl,--wrap=rand,--wrap=srand).
 *
 * The reference draws every uniform from one sequential glibc rand() stream
 * (photon.c:21,37,38,46).  glibc's generator is additive lagged-Fibonacci
 * (r[i] = r[i-3] + r[i-31]); with the walk's three to four draws per
 * absorption, its lag structure biases the reference's far-field heat by
 * about 1%.  Built against this file, the same reference object code draws
 * from SplitMix64 instead (Steele, Lea, Flood, "Fast splittable
 * pseudorandom number generators", OOPSLA 2014), whose outputs are returned
 * as 31-bit values in [0, RAND_MAX] like rand(). */
#include <stdint.h>

static uint64_t g_state = 0x853c49e6748fea9bull;

void __wrap_srand(unsigned int seed)
{
    g_state = 0x853c49e6748fea9bull ^ ((uint64_t)seed * 0x9E3779B97F4A7C15ull);
}

int __wrap_rand(void)
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (int)(z >> 33);
}
