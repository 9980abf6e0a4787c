/* The library's implicit per-process context and the parameter helpers
 * shared by the single-run, batched and serving entry points. */
#ifndef TMC_CONTEXT_H
#define TMC_CONTEXT_H

#include <hip/hip_runtime.h>

#include <functional>
#include <mutex>

#include "../../include/tiny_mc_gpu.h"
#include "tmc_plan.h"

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
