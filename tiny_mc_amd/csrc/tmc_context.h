/* The library's implicit per-process context and the parameter helpers
 * shared by the single-run, batched and serving entry points. */
#ifndef TMC_CONTEXT_H
#define TMC_CONTEXT_H

#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/tiny_mc_gpu.h"

namespace tmc {

struct Context {
    int device = -1;
    int compute_units = 0;
    char arch[64] = { 0 };
    hipStream_t stream = nullptr; /* default priority */
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    unsigned long long* d_tally = nullptr;
    size_t tally_words = 0;
    unsigned long long* d_counters = nullptr; /* 4 words */
};

/* Serialises every entry point that touches the context. */
std::mutex& api_mutex();

/* The context, created on the current device at first use.  Call with
 * api_mutex() held. */
int acquire_context(Context** out);

/* Parameter validation shared by every entry point (TMC_OK or
 * TMC_E_INVALID). */
int check_params(const tmc_params* p);

/* Whether the photon index range [first, first + photons) lies below 2^63:
 * then no photon index, chunk end or claim counter (which runs past the end
 * of the range by at most one chunk per wave) wraps around 64 bits. */
inline bool photon_range_ok(uint64_t first, uint64_t photons)
{
    const uint64_t limit = 1ull << 63;
    return first <= limit && photons <= limit - first;
}

/* photon.c:8: albedo = mu_s / (mu_s + mu_a) in binary32. */
float derived_albedo(const tmc_params* p);

/* photon.c:9: shells per mean free path through a double quotient. */
float derived_spm(const tmc_params* p);

/* Expected absorptions per photon of the FP32 weight cadence with the
 * roulette of photon.c:45-49, capped at kcap (scheduling estimate). */
double expected_interactions(const tmc_params* p, uint64_t kcap);

} // namespace tmc

#endif
