/* The no-allocation single run (tiny_mc_gpu_run_noalloc).
 *
 * Same walk as tiny_mc_gpu_run_ex (tmc::run_walk), but its device words (4
 * counters, then 4 * shells tally words) come from a static device array
 * instead of the context's allocation, so the call makes no hipMalloc,
 * hipFree or other allocation: for callers that cannot allocate.
 *
 * The array is a file-scope __device__ variable: every device the library
 * runs on holds its own copy until the process ends.  In host code its name
 * only stands for the variable; the device address is looked up once per
 * device, with that device current, when the context is created, and only
 * that address is ever passed to kernels. */
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/tiny_mc_gpu.h"
#include "tmc_context.h"

#define TMC_EXPORT extern "C" __attribute__((visibility("default")))

/* Counters, then the tallies of up to TMC_NOALLOC_MAX_SHELLS shells. */
static __device__ __align__(256) unsigned long long scratch[tmc::run_work_words(TMC_NOALLOC_MAX_SHELLS)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned long long* g_scratch[kMaxDevices] = {};

unsigned long long* scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch[device];
}

} // namespace

int tmc::scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return TMC_E_NODEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device])
        return TMC_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
        return TMC_E_HIP;
    g_scratch[device] = (unsigned long long*)ptr;
    return TMC_OK;
}

TMC_EXPORT int tiny_mc_gpu_run_noalloc(const tmc_params* p, uint64_t seed, const tmc_options* opt, uint64_t* tally,
                                       tmc_run_stats* stats)
{
    int rc = tmc::check_params(p);
    if (rc != TMC_OK)
        return rc;
    if (!tally)
        return TMC_E_INVALID;
    if (p->shells > TMC_NOALLOC_MAX_SHELLS)
        return TMC_E_TOO_LARGE;
    tmc_options o;
    if (opt)
        o = *opt;
    else
        tmc_options_default(&o);
    std::lock_guard<std::mutex> lock(tmc::api_mutex());
    tmc::Context* c = nullptr;
    rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    unsigned long long* work = scratch_of(c->device);
    if (!work) {
        /* the lookup at context creation failed: try once more, with the
         * context's device current (acquire_context made it so) */
        rc = tmc::scratch_setup(c->device);
        if (rc != TMC_OK)
            return rc;
        work = scratch_of(c->device);
    }
    return tmc::run_walk(c, p, seed, &o, work, tally, stats);
}
