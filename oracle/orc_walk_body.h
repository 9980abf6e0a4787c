/* Oracle walk template: one photon's random walk, restated from the reference.
 *
 * Test infrastructure only.  Included by oracle.c with these macros defined:
 *   W_FN                   name of the generated function
 *   W_REAL                 float (restatement) or double (FP64 oracle)
 *   W_SRC                  type of the uniform source
 *   W_SINK                 type of the tally sink
 *   W_U_STEP(s, k)         uniform in [0,1] for the step of interaction k
 *   W_U_DIR(s, k, a, i)    uniform i (0 or 1) of direction attempt a
 *   W_U_ROUL(s, k)         uniform for the roulette of interaction k
 *   W_LOG, W_SQRT          natural log and square root in W_REAL
 *   W_SINCOS2PI(a, s, c)   sin and cos of 2*pi*a (analytic sampler only)
 *   W_DEPOSIT(sink, shell, weight)  one absorption into a shell
 *
 * The body follows photon.c:6-51 step by step:
 *   launch            photon.c:11-18
 *   move              photon.c:21-24   t = -log(xi), x += t*u ...
 *   absorb            photon.c:26-32   shell index, clamp, deposit, weight *= albedo
 *   new direction     photon.c:34-43   rejection method (reference) or analytic
 *   roulette          photon.c:45-49   cutoff 0.001f, survival 0.1f, weight /= 0.1f
 * with the draws made in the same order as the reference (step, direction
 * attempts, roulette), so that the glibc-rand() instantiation consumes the
 * reference's stream exactly.  Deviations, both of probability zero under
 * the reference's own stream in practice:
 *   - the rejection loop also rejects t == 0 (the reference would divide 0/0);
 *   - the rejection loop gives up after ORC_MAX_ATTEMPTS attempts and keeps the
 *     last pair (probability 0.215^64 per interaction).
 * max_k truncates the walk after max_k interactions (0 = no limit); the
 * reference has no such limit.
 */
static uint64_t W_FN(W_SRC* src, W_SINK* sink, W_REAL albedo, W_REAL shells_per_mfp, uint32_t shells,
                     int sampler, uint64_t max_k)
{
    (void)src; /* the glibc rand() instantiation draws from the global stream */
    W_REAL x = 0, y = 0, z = 0;
    W_REAL u = 0, v = 0, w = 1;
    W_REAL weight = 1;
    uint64_t k = 0;

    for (;;) {
        W_REAL t = -W_LOG(W_U_STEP(src, k));
        x += t * u;
        y += t * v;
        z += t * w;

        unsigned int shell = W_SQRT(x * x + y * y + z * z) * shells_per_mfp;
        if (shell > shells - 1) {
            shell = shells - 1;
        }
        W_DEPOSIT(sink, shell, weight);
        weight *= albedo;

        if (sampler == ORC_SAMPLER_REJECTION) {
            W_REAL xi1, xi2;
            uint32_t a = 0;
            do {
                xi1 = (W_REAL)2 * W_U_DIR(src, k, a, 0) - (W_REAL)1;
                xi2 = (W_REAL)2 * W_U_DIR(src, k, a, 1) - (W_REAL)1;
                t = xi1 * xi1 + xi2 * xi2;
                ++a;
            } while (((W_REAL)1 < t || t == 0) && a < ORC_MAX_ATTEMPTS);
            u = (W_REAL)2 * t - (W_REAL)1;
            v = xi1 * W_SQRT(((W_REAL)1 - u * u) / t);
            w = xi2 * W_SQRT(((W_REAL)1 - u * u) / t);
        } else {
            const W_REAL mu = (W_REAL)2 * W_U_DIR(src, k, 0, 0) - (W_REAL)1;
            const W_REAL s = W_SQRT((W_REAL)1 - mu * mu);
            W_REAL sn, cs;
            W_SINCOS2PI(W_U_DIR(src, k, 0, 1), sn, cs);
            u = mu;
            v = s * cs;
            w = s * sn;
        }

        if (weight < (W_REAL)0.001f) {
            if (W_U_ROUL(src, k) > (W_REAL)0.1f) {
                ++k;
                break;
            }
            weight /= (W_REAL)0.1f;
        }
        ++k;
        if (max_k && k >= max_k) {
            break;
        }
    }
    return k;
}
