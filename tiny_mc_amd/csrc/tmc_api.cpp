/* C ABI of the photon-transport path (include/tiny_mc_gpu.h).
 *
 * Host side of one run: validate the params.h-style parameters, derive the
 * binary32 constants exactly as photon.c:8-9 does, clear the device tallies
 * and claim counter, launch the persistent walk kernel once, and copy the
 * 128-bit integer tallies back.  The conversion to the reference's heat /
 * heat2 units is done in double on the host.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

#include "../../include/tiny_mc_gpu.h"
#include "tmc_context.h"
#include "tmc_internal.h"

#define TMC_EXPORT extern "C" __attribute__((visibility("default")))

using tmc::Context;

namespace {

std::mutex g_mutex;
Context* g_ctx = nullptr;

int destroy_context(Context* c)
{
    if (!c)
        return TMC_OK;
    (void)hipSetDevice(c->device);
    if (c->d_work)
        tmc::release_device(c->d_work);
    if (c->ev_start)
        (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop)
        (void)hipEventDestroy(c->ev_stop);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
    return TMC_OK;
}

/* Creates the context on `device` (-1: the current device). */
int create_context(int device, Context** out)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return TMC_E_NODEVICE;
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess)
            return TMC_E_NODEVICE;
    }
    if (device >= count)
        return TMC_E_NODEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return TMC_E_HIP;
    Context* c = new (std::nothrow) Context;
    if (!c)
        return TMC_E_NOMEM;
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return TMC_E_HIP;
    }
    c->compute_units = prop.multiProcessorCount;
    std::snprintf(c->arch, sizeof c->arch, "%s", prop.gcnArchName);
    /* default-priority stream (never a prioritised one); no device memory
     * yet: the single-run words are allocated at the first run, so that the
     * no-allocation entry point can create the context too */
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess
        || hipEventCreate(&c->ev_start) != hipSuccess || hipEventCreate(&c->ev_stop) != hipSuccess) {
        destroy_context(c);
        return TMC_E_HIP;
    }
    /* the static scratch's address on this device; a failure here is
     * reported by tiny_mc_gpu_run_noalloc, the only user */
    (void)tmc::scratch_setup(device);
    *out = c;
    return TMC_OK;
}

} // namespace

std::mutex& tmc::api_mutex() { return g_mutex; }

int tmc::acquire_context(Context** out)
{
    if (!g_ctx) {
        int rc = create_context(-1, &g_ctx);
        if (rc != TMC_OK)
            return rc;
    }
    if (hipSetDevice(g_ctx->device) != hipSuccess)
        return TMC_E_HIP;
    *out = g_ctx;
    return TMC_OK;
}

namespace {

/* The single-run words at the largest shell count, once per context. */
int ensure_work(Context* c)
{
    if (c->d_work)
        return TMC_OK;
    if (hipMalloc((void**)&c->d_work, tmc::run_work_words(TMC_MAX_SHELLS) * sizeof(unsigned long long)) != hipSuccess) {
        c->d_work = nullptr;
        return TMC_E_NOMEM;
    }
    return TMC_OK;
}

/* Releases made while a serving kernel is resident (tmc_context.h). */
std::mutex g_release_mutex;
int g_serving_kernels = 0;
std::vector<std::function<void()>> g_deferred;

} // namespace

void tmc::release(std::function<void()> fn)
{
    std::lock_guard<std::mutex> lock(g_release_mutex);
    if (g_serving_kernels > 0)
        g_deferred.push_back(std::move(fn));
    else
        fn();
}

void tmc::release_device(void* p)
{
    if (p)
        release([p]() { (void)hipFree(p); });
}

void tmc::release_host(void* p)
{
    if (p)
        release([p]() { (void)hipHostFree(p); });
}

void tmc::serving_kernel_started()
{
    std::lock_guard<std::mutex> lock(g_release_mutex);
    ++g_serving_kernels;
}

void tmc::serving_kernel_left()
{
    std::lock_guard<std::mutex> lock(g_release_mutex);
    if (--g_serving_kernels > 0)
        return;
    std::vector<std::function<void()>> todo;
    todo.swap(g_deferred);
    for (auto& fn : todo)
        fn();
}

namespace {

double u128_to_double(uint64_t lo, uint64_t hi)
{
    /* hi * 2^64 + lo, with one rounding of the 128-bit value in the common
     * case hi == 0 */
    return std::ldexp((double)hi, 64) + (double)lo;
}

} // namespace

int tmc::run_walk(Context* c, const tmc_params* p, uint64_t seed, const tmc_options* o, unsigned long long* work,
                  uint64_t* tally, tmc_run_stats* st)
{
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t S = p->shells;
    if (o->sampler > TMC_SAMPLER_REJECTION || o->precision > TMC_PRECISION_FP64 || o->math > TMC_MATH_OWN)
        return TMC_E_INVALID;
    if (!tmc::photon_range_ok(o->first_photon, p->photons))
        return TMC_E_RANGE;

    uint32_t block = o->block ? o->block : 256u;
    if (block < 64u || block > 1024u || (block % 64u) != 0u)
        return TMC_E_INVALID;
    uint32_t chunk = o->chunk;
    if (chunk && (chunk < 64u || (chunk % 64u) != 0u || chunk > (1u << 24)))
        return TMC_E_INVALID;
    uint64_t kcap = o->max_interactions ? o->max_interactions : 0xFFFFFFFFull;
    if (kcap > 0xFFFFFFFFull)
        kcap = 0xFFFFFFFFull;
    if (o->reserved)
        return TMC_E_INVALID;
    const uint32_t table_k = tmc::weight_table_entries(p, o->precision, p->photons, o->weight_table);
    uint32_t grid = o->grid;
    if (grid == 0) {
        int per_cu = 0;
        if (tmc::walk_occupancy(o->sampler, o->precision, o->math, block, S, table_k, &per_cu) != hipSuccess
            || per_cu <= 0)
            return TMC_E_HIP;
        grid = (uint32_t)per_cu * (uint32_t)c->compute_units;
    }
    /* Do not launch far more waves than photons to walk. */
    const uint64_t waves_needed = (p->photons + 63u) / 64u;
    const uint32_t waves_per_wg = block / 64u;
    const uint64_t wg_needed = (waves_needed + waves_per_wg - 1) / waves_per_wg;
    if (wg_needed < grid)
        grid = (uint32_t)(wg_needed ? wg_needed : 1u);
    if (!chunk) {
        /* photons a wave claims at a time (64 per lane-photon): about 256
         * expected absorptions per lane between two claims, as a wave waits
         * for its claim's device-scope atomic (short walks would otherwise
         * claim every few steps, all waves on one counter), at least 4 per
         * lane, and few enough that each wave gets about 4 claims, so the
         * waves finish together */
        const double L = std::max(1.0, tmc::expected_interactions(p, kcap));
        const double waves = (double)grid * waves_per_wg;
        double per_lane = std::ceil(256.0 / L);
        per_lane = std::min(per_lane, std::floor((double)p->photons / (64.0 * 4.0 * waves)));
        per_lane = std::min(std::max(per_lane, 4.0), (double)(1u << 18));
        chunk = 64u * (uint32_t)per_lane;
    }

    unsigned long long* const d_counters = work;
    unsigned long long* const d_tally = work + tmc::kRunCounterWords;
    tmc::WalkArgs a;
    std::memset(&a, 0, sizeof a);
    a.seed = seed;
    a.first_photon = o->first_photon;
    a.photons = p->photons;
    a.albedo = tmc::derived_albedo(p);
    a.spm = tmc::derived_spm(p);
    a.shell_max = (float)(S - 1);
    a.shells = S;
    a.chunk = chunk;
    /* precision variants: constants in FP64, each rounded once */
    const double mu_a = p->mu_a, mu_s = p->mu_s;
    a.albedo_d = mu_s / (mu_s + mu_a);
    a.spm_d = 1e4 / (double)p->microns_per_shell / (mu_a + mu_s);
    const double l2a = std::log2(a.albedo_d), l2s = -std::log2((double)0.1f);
    a.l2a_hi = (float)l2a;
    a.l2a_lo = (float)(l2a - (double)a.l2a_hi);
    a.l2s_hi = (float)l2s;
    a.l2s_lo = (float)(l2s - (double)a.l2s_hi);
    a.spm_w = (float)a.spm_d;
    /* Each wave swaps the workgroup's LDS sums out at least every
     * flush_interval steps, so between two swaps of one entry at most
     * waves_per_wg * 64 * flush_interval deposits of at most 2^43 (heat2
     * image of w = 1) accumulate: keep that below 2^63. */
    a.flush_interval = (1u << 14) / waves_per_wg;
    a.kcap = (uint32_t)kcap;
    a.table_k = table_k;
    a.counters = d_counters;
    a.tally = d_tally;

    float kernel_ms = 0.0f;
    /* a fill kernel, not a memset: the same sequence serves callers that
     * cannot have memset nodes in their streams */
    if (tmc::launch_clear(work, tmc::run_work_words(S), c->stream) != hipSuccess)
        return TMC_E_HIP;
    unsigned long long counters[4] = { 0, 0, 0, 0 };
    if (p->photons > 0) {
        if (hipEventRecord(c->ev_start, c->stream) != hipSuccess)
            return TMC_E_HIP;
        if (tmc::launch_walk(o->sampler, o->precision, o->math, a, grid, block, c->stream) != hipSuccess)
            return TMC_E_HIP;
        if (hipEventRecord(c->ev_stop, c->stream) != hipSuccess)
            return TMC_E_HIP;
    }
    if (hipMemcpyAsync(tally, d_tally, 4 * (size_t)S * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream)
            != hipSuccess
        || hipMemcpyAsync(counters, d_counters, sizeof counters, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return TMC_E_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return TMC_E_HIP;
    if (p->photons > 0 && hipEventElapsedTime(&kernel_ms, c->ev_start, c->ev_stop) != hipSuccess)
        return TMC_E_HIP;
    if (st) {
        st->photons = p->photons;
        st->interactions = counters[1];
        st->truncated = counters[2];
        st->kernel_ms = kernel_ms;
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st->grid = grid;
        st->block = block;
        st->chunk = chunk;
        st->sampler = o->sampler;
        st->precision = o->precision;
        st->math = o->math;
        st->weight_table = table_k;
    }
    return TMC_OK;
}

TMC_EXPORT void tmc_options_default(tmc_options* opt)
{
    if (!opt)
        return;
    std::memset(opt, 0, sizeof *opt);
    opt->sampler = TMC_SAMPLER_ANALYTIC;
}

TMC_EXPORT int tiny_mc_gpu_run_ex(const tmc_params* p, uint64_t seed, const tmc_options* opt, uint64_t* tally,
                                  tmc_run_stats* stats)
{
    int rc = tmc::check_params(p);
    if (rc != TMC_OK)
        return rc;
    if (!tally)
        return TMC_E_INVALID;
    tmc_options o;
    if (opt)
        o = *opt;
    else
        tmc_options_default(&o);
    std::lock_guard<std::mutex> lock(g_mutex);
    Context* c = nullptr;
    rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    rc = ensure_work(c);
    if (rc != TMC_OK)
        return rc;
    return tmc::run_walk(c, p, seed, &o, c->d_work, tally, stats);
}

TMC_EXPORT int tiny_mc_gpu_tally_to_heat(const tmc_params* p, const uint64_t* tally, double* heat, double* heat2)
{
    int rc = tmc::check_params(p);
    if (rc != TMC_OK)
        return rc;
    if (!tally || !heat || !heat2)
        return TMC_E_INVALID;
    const uint32_t S = p->shells;
    /* (1 - albedo) in double from the parameters; the FP32 transport's
     * albedo (photon.c:8) only drives the weight decay. */
    const double albedo = (double)p->mu_s / ((double)p->mu_s + (double)p->mu_a);
    const double oma = 1.0 - albedo;
    const double s1 = std::ldexp(oma, -TMC_HEAT_FRAC_BITS);
    const double s2 = std::ldexp(oma * oma, -TMC_HEAT2_FRAC_BITS);
    for (uint32_t s = 0; s < S; ++s) {
        heat[s] = s1 * u128_to_double(tally[2 * s], tally[2 * s + 1]);
        heat2[s] = s2 * u128_to_double(tally[2 * S + 2 * s], tally[2 * S + 2 * s + 1]);
    }
    return TMC_OK;
}

TMC_EXPORT int tiny_mc_gpu_run(const tmc_params* p, uint64_t seed, double* heat, double* heat2)
{
    int rc = tmc::check_params(p);
    if (rc != TMC_OK)
        return rc;
    if (!heat || !heat2)
        return TMC_E_INVALID;
    uint64_t* tally = new (std::nothrow) uint64_t[4 * (size_t)p->shells];
    if (!tally)
        return TMC_E_NOMEM;
    rc = tiny_mc_gpu_run_ex(p, seed, nullptr, tally, nullptr);
    if (rc == TMC_OK)
        rc = tiny_mc_gpu_tally_to_heat(p, tally, heat, heat2);
    delete[] tally;
    return rc;
}

TMC_EXPORT int tiny_mc_gpu_init(int device)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ctx) {
        if (device >= 0 && device != g_ctx->device)
            return TMC_E_INVALID;
        return TMC_OK;
    }
    return create_context(device, &g_ctx);
}

TMC_EXPORT void tiny_mc_gpu_shutdown(void)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    destroy_context(g_ctx);
    g_ctx = nullptr;
}

TMC_EXPORT int tiny_mc_gpu_device_info(char* name, size_t cap, int* compute_units)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    Context* c = nullptr;
    int rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    if (name && cap)
        std::snprintf(name, cap, "%s", c->arch);
    if (compute_units)
        *compute_units = c->compute_units;
    return TMC_OK;
}

TMC_EXPORT double tiny_mc_gpu_expected_interactions(const tmc_params* p)
{
    if (tmc::check_params(p) != TMC_OK)
        return 0.0;
    return tmc::expected_interactions(p, 0);
}

TMC_EXPORT const char* tiny_mc_gpu_strerror(int status)
{
    switch (status) {
    case TMC_OK: return "ok";
    case TMC_E_INVALID: return "invalid argument";
    case TMC_E_NODEVICE: return "no usable HIP device";
    case TMC_E_HIP: return "HIP runtime error";
    case TMC_E_NOMEM: return "out of memory";
    case TMC_E_RANGE: return "photon index range exceeds 2^63";
    case TMC_E_UNSUPPORTED: return "unsupported option";
    case TMC_E_TOO_LARGE: return "more shells than the no-allocation scratch holds";
    case TMC_PENDING: return "pending";
    default: return "unknown status";
    }
}

namespace {

/* Device buffers of a diagnostic call, freed on every path. */
struct DevBufs {
    void* p[5] = { nullptr, nullptr, nullptr, nullptr, nullptr };
    ~DevBufs()
    {
        for (void* q : p)
            tmc::release_device(q);
    }
};

} // namespace

TMC_EXPORT int tiny_mc_gpu_philox(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out)
{
    if ((!ctr || !key || !out) && n)
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lock(g_mutex);
    Context* c = nullptr;
    int rc = tmc::acquire_context(&c);
    if (rc != TMC_OK || n == 0)
        return rc;
    DevBufs b;
    if (hipMalloc(&b.p[0], 16 * (size_t)n) != hipSuccess || hipMalloc(&b.p[1], 8 * (size_t)n) != hipSuccess
        || hipMalloc(&b.p[2], 16 * (size_t)n) != hipSuccess)
        return TMC_E_NOMEM;
    if (hipMemcpyAsync(b.p[0], ctr, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess
        || hipMemcpyAsync(b.p[1], key, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess
        || tmc::launch_philox((const uint32_t*)b.p[0], (const uint32_t*)b.p[1], n, (uint32_t*)b.p[2], c->stream)
            != hipSuccess
        || hipMemcpyAsync(out, b.p[2], 16 * (size_t)n, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipStreamSynchronize(c->stream) != hipSuccess)
        return TMC_E_HIP;
    return TMC_OK;
}

TMC_EXPORT int tiny_mc_gpu_eval_step_math(uint32_t math, const float* in, uint32_t n, float* step_out,
                                          float* sin_out, float* cos_out, float* polar_out)
{
    if (((!in || !step_out || !sin_out || !cos_out || !polar_out) && n) || math > TMC_MATH_OWN)
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lock(g_mutex);
    Context* c = nullptr;
    int rc = tmc::acquire_context(&c);
    if (rc != TMC_OK || n == 0)
        return rc;
    DevBufs b;
    const size_t bytes = 4 * (size_t)n;
    for (int i = 0; i < 5; ++i)
        if (hipMalloc(&b.p[i], bytes) != hipSuccess)
            return TMC_E_NOMEM;
    if (hipMemcpyAsync(b.p[0], in, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess
        || tmc::launch_eval_step_math(math, (const float*)b.p[0], n, (float*)b.p[1], (float*)b.p[2],
                                      (float*)b.p[3], (float*)b.p[4], c->stream)
            != hipSuccess
        || hipMemcpyAsync(step_out, b.p[1], bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipMemcpyAsync(sin_out, b.p[2], bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipMemcpyAsync(cos_out, b.p[3], bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipMemcpyAsync(polar_out, b.p[4], bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipStreamSynchronize(c->stream) != hipSuccess)
        return TMC_E_HIP;
    return TMC_OK;
}

TMC_EXPORT int tiny_mc_gpu_eval_math(const float* in, uint32_t n, float* log_out, float* sin_out, float* cos_out)
{
    if ((!in || !log_out || !sin_out || !cos_out) && n)
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lock(g_mutex);
    Context* c = nullptr;
    int rc = tmc::acquire_context(&c);
    if (rc != TMC_OK || n == 0)
        return rc;
    DevBufs b;
    const size_t bytes = 4 * (size_t)n;
    for (int i = 0; i < 4; ++i)
        if (hipMalloc(&b.p[i], bytes) != hipSuccess)
            return TMC_E_NOMEM;
    if (hipMemcpyAsync(b.p[0], in, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess
        || tmc::launch_eval_math((const float*)b.p[0], n, (float*)b.p[1], (float*)b.p[2], (float*)b.p[3], c->stream)
            != hipSuccess
        || hipMemcpyAsync(log_out, b.p[1], bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipMemcpyAsync(sin_out, b.p[2], bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipMemcpyAsync(cos_out, b.p[3], bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipStreamSynchronize(c->stream) != hipSuccess)
        return TMC_E_HIP;
    return TMC_OK;
}

TMC_EXPORT int tiny_mc_gpu_check_short_op(uint32_t op, uint32_t first_bits, uint64_t count, uint64_t* mismatches,
                                          uint32_t* first_mismatch)
{
    if (!mismatches || !first_mismatch || op > TMC_CHECK_DIV_TENTH || count > (1ull << 32)
        || (uint64_t)first_bits + count > (1ull << 32))
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lock(g_mutex);
    Context* c = nullptr;
    int rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    DevBufs b;
    unsigned long long h[2] = { 0ull, ~0ull };
    if (hipMalloc(&b.p[0], sizeof h) != hipSuccess)
        return TMC_E_NOMEM;
    if (hipMemcpyAsync(b.p[0], h, sizeof h, hipMemcpyHostToDevice, c->stream) != hipSuccess
        || tmc::launch_check_short_op(op, first_bits, count, (unsigned long long*)b.p[0], c->stream) != hipSuccess
        || hipMemcpyAsync(h, b.p[0], sizeof h, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipStreamSynchronize(c->stream) != hipSuccess)
        return TMC_E_HIP;
    *mismatches = h[0];
    *first_mismatch = h[0] ? (uint32_t)h[1] : 0u;
    return TMC_OK;
}

TMC_EXPORT int tiny_mc_gpu_check_shell_index(const tmc_params* p, uint32_t first_bits, uint64_t count,
                                             uint64_t* violations, uint32_t* first_violation, uint64_t* unsure)
{
    if (!violations || !first_violation || !unsure || count > (1ull << 32)
        || (uint64_t)first_bits + count > (1ull << 32))
        return TMC_E_INVALID;
    int rc = tmc::check_params(p);
    if (rc != TMC_OK)
        return rc;
    std::lock_guard<std::mutex> lock(g_mutex);
    Context* c = nullptr;
    rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    DevBufs b;
    unsigned long long h[3] = { 0ull, ~0ull, 0ull };
    if (hipMalloc(&b.p[0], sizeof h) != hipSuccess)
        return TMC_E_NOMEM;
    if (hipMemcpyAsync(b.p[0], h, sizeof h, hipMemcpyHostToDevice, c->stream) != hipSuccess
        || tmc::launch_check_shell(tmc::derived_spm(p), (float)(p->shells - 1), first_bits, count,
                                   (unsigned long long*)b.p[0], c->stream)
            != hipSuccess
        || hipMemcpyAsync(h, b.p[0], sizeof h, hipMemcpyDeviceToHost, c->stream) != hipSuccess
        || hipStreamSynchronize(c->stream) != hipSuccess)
        return TMC_E_HIP;
    *violations = h[0];
    *first_violation = h[0] ? (uint32_t)h[1] : 0u;
    *unsure = h[2];
    return TMC_OK;
}

TMC_EXPORT int tiny_mc_gpu_check_sqrt(uint32_t first_bits, uint64_t count, uint64_t* mismatches,
                                      uint32_t* first_mismatch)
{
    return tiny_mc_gpu_check_short_op(TMC_CHECK_SQRT, first_bits, count, mismatches, first_mismatch);
}
