/* The library's implicit per-process context and the parameter helpers
 * shared by the single-run, batched and serving entry points. */
#ifndef TMC_CONTEXT_H
#define TMC_CONTEXT_H

#include <hip/hip_runtime.h>

#include <functional>
#include <mutex>

#include "../../include/tiny_mc_gpu.h"

namespace tmc {

/* Device words of one single run: the walk's counters ([0] claim offset,
 * [1] absorptions, [2] truncated photons, [3] unused), then the 4 * shells
 * tally words.  The context allocates them once at the largest shell count
 * (run_work_words(TMC_MAX_SHELLS) words, 256 KiB) and keeps them until
 * shutdown, so no single run allocates or frees device memory. */
constexpr size_t kRunCounterWords = 4;
constexpr size_t run_work_words(uint32_t shells) { return kRunCounterWords + 4 * (size_t)shells; }

struct Context {
    int device = -1;
    int compute_units = 0;
    char arch[64] = { 0 };
    hipStream_t stream = nullptr; /* default priority */
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    unsigned long long* d_work = nullptr; /* run_work_words(TMC_MAX_SHELLS) words, allocated at the first run */
};

/* Serialises every entry point that touches the context. */
std::mutex& api_mutex();

/* The context, created on the current device at first use.  Call with
 * api_mutex() held. */
int acquire_context(Context** out);

/* One single run (tiny_mc_gpu_run_ex) with its counters and tallies in the
 * device words `work` (run_work_words(p->shells) of them): clears them with
 * a kernel, launches the walk, copies the tallies to `tally` (host) and
 * synchronises the context's stream.  Allocates nothing.  Call with
 * api_mutex() held; p is checked. */
int run_walk(Context* c, const tmc_params* p, uint64_t seed, const tmc_options* o, unsigned long long* work,
             uint64_t* tally, tmc_run_stats* st);

/* The static scratch of tiny_mc_gpu_run_noalloc on `device` (current):
 * looks its address up once per device.  TMC_OK or a negative status. */
int scratch_setup(int device);

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

/* The binary32 weight cadence: absorptions k1 of the first cycle and k2 of
 * the second (false if the first is longer than `limit`). */
bool weight_cadence(const tmc_params* p, uint32_t limit, uint32_t* k1, uint32_t* k2);

/* Entries of the single run's weight table for these parameters and the
 * caller's request (tmc_options.weight_table; 0: none). */
uint32_t weight_table_entries(const tmc_params* p, uint32_t precision, uint32_t sampler, uint64_t photons,
                              uint32_t request);

/* Releases of device resources while a serving kernel may be resident.
 * hipFree and hipHostFree synchronise the whole device, so under load they
 * would wait for a resident serving kernel that never becomes idle.  A
 * release made while any server's kernel runs is queued and performed once
 * no serving kernel runs; otherwise it happens at once.  A server calls
 * serving_kernel_started() before it launches its kernel (which waits for
 * a release in progress) and serving_kernel_left() once the kernel has
 * ended (which performs the queued releases when it was the last one). */
void release(std::function<void()> fn);
void release_device(void* p); /* hipFree */
void release_host(void* p); /* hipHostFree */
void serving_kernel_started();
void serving_kernel_left();

} // namespace tmc

#endif
