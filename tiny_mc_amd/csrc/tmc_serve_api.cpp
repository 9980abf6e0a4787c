/* Serving: continuous batching of simulation requests on a resident
 * persistent kernel (include/tiny_mc_gpu.h, tmc_serve_*).
 *
 * Memory the server keeps for its lifetime:
 *   device          tally arena (128-bit sums of every resident request),
 *                   per-slot counters (photons completed, absorptions), the
 *                   ticket counter and a cache of the control words;
 *   host-coherent   the request table (one BatchSet per slot), the chunk
 *                   ring, the control words (tail, stop), a mirror of the
 *                   tally arena that receives finished sums, and per-slot
 *                   flags.
 * The server thread makes no copies and waits on no queue while the kernel
 * runs: it writes descriptors, chunks and the tail with ordinary stores and
 * reads finished requests from the mirror once their flag holds their
 * token.  (A copy could sit behind another process's or thread's kernel
 * that in turn waits for the compute units the resident kernel holds.)
 *
 * The server thread loops:
 *   1. notice a kernel that has left (it leaves only when asked to stop);
 *   2. admit queued requests in arrival order while a slot and room in the
 *      tally arena are free;
 *   3. publish chunks, shortest remaining work first, keeping about one
 *      published-but-unclaimed chunk per resident wave, so a request that
 *      arrives while others run gets its first chunks claimed within about
 *      one chunk time of a wave (this is what folds it into the running
 *      kernel);
 *   4. launch the kernel when published work waits and none is running;
 *   5. finish every request whose flag holds its token: take its sums from
 *      the mirror, record its latency, free its slot and arena region.
 * After idle_ms with no request in the server it asks the kernel to stop:
 * then every published chunk has been claimed and done, so each wave's
 * ticket lies past the tail and the next launch starts its tickets at the
 * tail.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/tiny_mc_gpu.h"
#include "tmc_batch_internal.h"
#include "tmc_context.h"

#define TMC_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    tmc_set set;
    uint32_t slot = 0, gen = 0;
    uint64_t tally_off = 0, words = 0;
    uint64_t chunk = 0; /* photons per chunk */
    uint64_t next = 0; /* next unpublished photon offset */
    double L = 1.0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t interactions = 0;
    std::vector<uint64_t> result;
};

/* First-fit allocator of the tally arena (units: 64-bit words). */
struct Arena {
    std::map<uint64_t, uint64_t> free_; /* offset -> size */
    void init(uint64_t words) { free_ = { { 0, words } }; }
    bool alloc(uint64_t n, uint64_t* off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest)
                    free_[at] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n)
    {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

/* Workgroup slots the serving kernel leaves free, so other work on the
 * device (another thread's run, the runtime's copies) still gets compute
 * units while requests keep the kernel resident. */
constexpr uint32_t kFreeSlots = 8;
/* Minimum time between two reads of the host's control words by the
 * kernel, in 100 MHz realtime ticks (2 us). */
constexpr uint32_t kRefreshTicks = 200;

} // namespace

struct tmc_server {
    tmc_serve_options opt;
    int device = 0;
    uint32_t grid = 0, block = 0, total_waves = 0;
    hipStream_t kstream = nullptr;
    /* device */
    unsigned long long* d_tally = nullptr;
    unsigned long long* d_done = nullptr;
    unsigned long long* d_counters = nullptr;
    tmc::ServeCache* d_cache = nullptr;
    /* host-coherent */
    tmc::ServeControl* h_ctl = nullptr;
    tmc::BatchSet* h_sets = nullptr;
    tmc::BatchChunk* h_ring = nullptr;
    unsigned long long* h_out = nullptr;
    unsigned long long* h_flags = nullptr;
    tmc::BatchArgs args;

    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool closing = false;
    int error = TMC_OK;
    Clock::time_point t0;

    uint64_t next_id = 1;
    std::deque<Request*> pending;
    std::vector<Request*> active; /* admitted, unfinished */
    std::map<uint64_t, Request*> finished;
    std::map<uint64_t, int> state; /* id -> 0 queued/active, 1 finished */
    std::vector<uint32_t> free_slots;
    std::vector<uint32_t> slot_gen;
    Arena arena;
    uint64_t published = 0; /* ring indices written (the tail) */
    uint64_t claimed = 0; /* highest claim the kernel reported */
    uint64_t base = 0; /* first ticket of the next launch */
    uint64_t launch_id = 0;
    bool kernel_running = false;
    bool stopping = false; /* the running kernel was asked to stop */
    double last_busy = 0; /* last time the server held a request */
    tmc_serve_stats st;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void server_free(tmc_server* s)
{
    if (s->kstream)
        (void)hipStreamSynchronize(s->kstream);
    for (void* d : { (void*)s->d_tally, (void*)s->d_done, (void*)s->d_counters, (void*)s->d_cache })
        if (d)
            (void)hipFree(d);
    for (void* h : { (void*)s->h_ctl, (void*)s->h_sets, (void*)s->h_ring, (void*)s->h_out, (void*)s->h_flags })
        if (h)
            (void)hipHostFree(h);
    if (s->kstream)
        (void)hipStreamDestroy(s->kstream);
    for (Request* r : s->pending)
        delete r;
    for (Request* r : s->active)
        delete r;
    for (auto& kv : s->finished)
        delete kv.second;
    delete s;
}

/* Chunk size of a request: about chunk_work absorptions (so a chunk lasts
 * about chunk_work / 64 steps of a wave, which bounds how long a new
 * request waits for waves), between 1 and max_photons_per_lane photons per
 * lane, never more than the request. */
uint64_t chunk_photons(const tmc_serve_options& o, double L, uint64_t N)
{
    double per_lane = std::floor((double)o.chunk_work / (64.0 * L));
    per_lane = std::min(std::max(per_lane, 1.0), (double)o.max_photons_per_lane);
    return std::max<uint64_t>(1, std::min<uint64_t>(N, 64u * (uint64_t)per_lane));
}

void store_release(uint64_t* p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
uint64_t load_acquire(const uint64_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }

/* Asks the running kernel to stop; called only when no request is admitted
 * or queued and nothing will be published until it has left. */
void request_stop(tmc_server* s)
{
    if (s->kernel_running && !s->stopping) {
        store_release(&s->h_ctl->stop, s->launch_id);
        s->stopping = true;
    }
}

/* Notices a kernel that has left.  It leaves only after a stop request,
 * when every published chunk had been claimed, so the next launch starts
 * its tickets at the tail. */
void reap_kernel(tmc_server* s)
{
    if (!s->kernel_running)
        return;
    const hipError_t q = hipStreamQuery(s->kstream);
    if (q == hipErrorNotReady)
        return;
    s->kernel_running = false;
    s->stopping = false;
    if (q != hipSuccess) {
        s->error = TMC_E_HIP;
        return;
    }
    s->base = s->published;
    s->claimed = std::max(s->claimed, s->published);
}

/* One pass of the server loop; called with s->m held. */
void serve_step(tmc_server* s)
{
    const uint64_t ring = s->opt.ring_chunks;
    reap_kernel(s);
    if (s->error != TMC_OK)
        return;

    /* admit in arrival order.  Tally regions and counters are zero here:
     * zeroed at open, and again by the wave that finished the region's
     * previous request. */
    while (!s->pending.empty() && !s->free_slots.empty()) {
        Request* r = s->pending.front();
        uint64_t off = 0;
        if (!s->arena.alloc(r->words, &off))
            break;
        s->pending.pop_front();
        r->slot = s->free_slots.back();
        s->free_slots.pop_back();
        r->gen = (s->slot_gen[r->slot] + 1) & 0xFFFFu;
        s->slot_gen[r->slot] = r->gen;
        r->tally_off = off;
        r->t_admit = s->now();
        const tmc_params& q = r->set.params;
        tmc::BatchSet b;
        std::memset(&b, 0, sizeof b);
        b.seed = r->set.seed;
        b.first_photon = r->set.first_photon;
        b.tally_offset = off;
        b.albedo = tmc::derived_albedo(&q);
        b.spm = tmc::derived_spm(&q);
        b.shell_max = (float)(q.shells - 1);
        b.shells = q.shells;
        b.kcap = 0xFFFFFFFFu;
        b.photons = q.photons;
        b.token = r->id;
        /* read by the kernel only after a chunk of this request is
         * published, which the release store of the tail orders after */
        std::memcpy(&s->h_sets[r->slot], &b, sizeof b);
        s->active.push_back(r);
        s->st.admitted++;
    }

    /* publish, shortest remaining work first, keeping about one unclaimed
     * chunk per resident wave.  A ring index is rewritten only after the
     * tail has moved more than ring - 2 * waves past the highest claim
     * reported.  A wave reads its entry at the first poll after the tail
     * passes its ticket (within 16 of its steps), while rewriting an entry
     * it has not read would take publishing about a million chunks first. */
    s->claimed = std::max<uint64_t>(s->claimed, load_acquire(&s->h_ctl->claimed));
    const uint64_t backlog_max = std::min<uint64_t>(ring - 2 * s->total_waves, s->total_waves);
    auto unclaimed = [s]() -> uint64_t { return s->published > s->claimed ? s->published - s->claimed : 0; };
    const uint64_t before = s->published;
    if (!(s->kernel_running && s->stopping) && unclaimed() < backlog_max) {
        std::vector<Request*> order;
        for (Request* r : s->active)
            if (r->next < r->set.params.photons)
                order.push_back(r);
        std::sort(order.begin(), order.end(), [](const Request* x, const Request* y) {
            return (double)(x->set.params.photons - x->next) * x->L < (double)(y->set.params.photons - y->next) * y->L;
        });
        for (Request* r : order) {
            const uint64_t N = r->set.params.photons;
            while (r->next < N && unclaimed() < backlog_max) {
                const uint64_t cnt = std::min(r->chunk, N - r->next);
                tmc::BatchChunk& c = s->h_ring[s->published & (ring - 1)];
                c.set = r->slot | (r->gen << TMC_SERVE_SLOT_BITS);
                c.count = (uint32_t)cnt;
                c.first = r->next;
                r->next += cnt;
                s->published++;
            }
            if (unclaimed() >= backlog_max)
                break;
        }
    }
    if (s->published != before) {
        store_release(&s->h_ctl->tail, s->published);
        s->st.chunks_published = s->published;
    }

    /* keep a kernel running while published work waits */
    if (!s->kernel_running && s->published > s->base) {
        s->args.ticket_base = s->base;
        s->args.launch_id = ++s->launch_id;
        if (tmc::launch_serve(s->opt.sampler, s->opt.math, s->args, s->grid, s->block, s->kstream) != hipSuccess) {
            s->error = TMC_E_HIP;
            return;
        }
        s->kernel_running = true;
        s->st.kernel_launches++;
    }

    /* finished requests: the flag holds the token once the sums are in the
     * mirror */
    const double t = s->now();
    bool any = false;
    for (size_t i = 0; i < s->active.size();) {
        Request* r = s->active[i];
        if (r->next < r->set.params.photons || load_acquire((const uint64_t*)s->h_flags + 2 * r->slot) != r->id) {
            ++i;
            continue;
        }
        const unsigned long long* src = s->h_out + r->tally_off;
        r->result.assign(src, src + r->words);
        r->interactions = s->h_flags[2 * (uint64_t)r->slot + 1];
        r->t_done = t;
        s->arena.release(r->tally_off, r->words);
        s->free_slots.push_back(r->slot);
        s->active[i] = s->active.back();
        s->active.pop_back();
        s->finished[r->id] = r;
        s->state[r->id] = 1;
        s->st.completed++;
        s->st.photons_completed += r->set.params.photons;
        s->st.interactions_completed += r->interactions;
        any = true;
    }
    if (any)
        s->cv_done.notify_all();
}

void serve_loop(tmc_server* s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->m);
    for (;;) {
        const bool idle = s->pending.empty() && s->active.empty();
        if (idle && s->closing)
            break;
        if (idle) {
            s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            if (s->now() - s->last_busy > 1e-3 * s->opt.idle_ms)
                request_stop(s);
            reap_kernel(s);
            if (s->error != TMC_OK)
                break;
            continue;
        }
        s->last_busy = s->now();
        serve_step(s);
        if (s->error != TMC_OK)
            break;
        /* short pause: other threads submit and poll under the lock */
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
    }
    s->cv_done.notify_all();
    /* every request is finished (or the server failed): let the kernel go */
    if (s->kernel_running) {
        store_release(&s->h_ctl->stop, s->launch_id);
        lk.unlock();
        (void)hipStreamSynchronize(s->kstream);
        lk.lock();
        s->kernel_running = false;
    }
}

template <typename T>
bool host_alloc(T** p, size_t count)
{
    if (hipHostMalloc((void**)p, count * sizeof(T), hipHostMallocCoherent) != hipSuccess)
        return false;
    std::memset((void*)*p, 0, count * sizeof(T));
    return true;
}

template <typename T>
T* device_view(T* host)
{
    void* d = nullptr;
    return hipHostGetDevicePointer(&d, host, 0) == hipSuccess ? (T*)d : nullptr;
}

} // namespace

TMC_EXPORT void tmc_serve_options_default(tmc_serve_options* opt)
{
    if (opt)
        std::memset(opt, 0, sizeof *opt);
}

TMC_EXPORT int tmc_serve_open(const tmc_serve_options* opt, tmc_server** out)
{
    if (!out)
        return TMC_E_INVALID;
    *out = nullptr;
    tmc_serve_options o;
    if (opt)
        o = *opt;
    else
        tmc_serve_options_default(&o);
    if (o.sampler > TMC_SAMPLER_REJECTION || o.math > TMC_MATH_OWN)
        return TMC_E_INVALID;
    if (!o.max_shells)
        o.max_shells = 401;
    if (!o.max_requests)
        o.max_requests = 4096;
    if (!o.block)
        o.block = 256;
    if (!o.ring_chunks)
        o.ring_chunks = 1ull << 20;
    if (!o.chunk_work)
        o.chunk_work = 1ull << 17;
    if (!o.max_photons_per_lane)
        o.max_photons_per_lane = 16;
    if (!o.idle_ms)
        o.idle_ms = 50;
    if (!o.tally_words)
        o.tally_words = (uint64_t)o.max_requests * 4 * o.max_shells;
    /* a chunk (at most 64 * max_photons_per_lane photons) keeps its photon
     * count in 32 bits */
    if (o.max_shells > TMC_MAX_SHELLS || o.max_requests > (1u << TMC_SERVE_SLOT_BITS) || o.block % 64
        || o.block > 1024 || (o.ring_chunks & (o.ring_chunks - 1)) || o.ring_chunks > (1ull << 31)
        || o.max_photons_per_lane > (1u << 24))
        return TMC_E_INVALID;

    std::lock_guard<std::mutex> lock(tmc::api_mutex());
    tmc::Context* c = nullptr;
    int rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    tmc_server* s = new (std::nothrow) tmc_server;
    if (!s)
        return TMC_E_NOMEM;
    s->opt = o;
    s->device = c->device;
    std::memset(&s->st, 0, sizeof s->st);
    int wg = 0;
    if (tmc::batch_lds_bytes(o.block, o.max_shells, 1) > 160u * 1024u
        || tmc::serve_occupancy(o.sampler, o.math, o.block, o.max_shells, 1, &wg) != hipSuccess || wg <= 0) {
        delete s;
        return TMC_E_INVALID;
    }
    s->block = o.block;
    s->grid = (uint32_t)wg * (uint32_t)c->compute_units;
    s->grid = s->grid > 2 * kFreeSlots ? s->grid - kFreeSlots : (s->grid + 1) / 2;
    s->total_waves = s->grid * (o.block / 64);
    if (o.ring_chunks <= 2ull * s->total_waves) {
        delete s;
        return TMC_E_INVALID;
    }
    const size_t R = o.max_requests;
    if (hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking) != hipSuccess
        || hipMalloc((void**)&s->d_tally, o.tally_words * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_done, 2 * R * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_counters, 4 * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_cache, sizeof(tmc::ServeCache)) != hipSuccess
        || !host_alloc(&s->h_ctl, 1) || !host_alloc(&s->h_sets, R) || !host_alloc(&s->h_ring, o.ring_chunks)
        || !host_alloc(&s->h_out, o.tally_words) || !host_alloc(&s->h_flags, 2 * R)) {
        server_free(s);
        return TMC_E_NOMEM;
    }
    if (hipMemset(s->d_tally, 0, o.tally_words * sizeof(uint64_t)) != hipSuccess
        || hipMemset(s->d_done, 0, 2 * R * sizeof(uint64_t)) != hipSuccess
        || hipMemset(s->d_counters, 0, 4 * sizeof(uint64_t)) != hipSuccess
        || hipMemset(s->d_cache, 0, sizeof(tmc::ServeCache)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        server_free(s);
        return TMC_E_HIP;
    }
    s->arena.init(o.tally_words);
    s->slot_gen.assign(R, 0);
    for (uint32_t i = 0; i < R; ++i)
        s->free_slots.push_back((uint32_t)(R - 1 - i));

    std::memset(&s->args, 0, sizeof s->args);
    s->args.sets = device_view(s->h_sets);
    s->args.chunks = device_view(s->h_ring);
    s->args.hctl = device_view(s->h_ctl);
    s->args.out = device_view(s->h_out);
    s->args.flags = device_view(s->h_flags);
    if (!s->args.sets || !s->args.chunks || !s->args.hctl || !s->args.out || !s->args.flags) {
        server_free(s);
        return TMC_E_HIP;
    }
    s->args.nchunks = 0;
    s->args.shells_max = o.max_shells;
    s->args.copies = 1;
    s->args.ring_mask = (uint32_t)(o.ring_chunks - 1);
    s->args.counters = s->d_counters;
    s->args.tally = s->d_tally;
    s->args.done = s->d_done;
    s->args.dctl = s->d_cache;
    s->args.total_waves = s->total_waves;
    s->args.refresh_ticks = kRefreshTicks;

    s->t0 = Clock::now();
    s->worker = std::thread(serve_loop, s);
    *out = s;
    return TMC_OK;
}

TMC_EXPORT int tmc_serve_submit(tmc_server* s, const tmc_set* req, uint64_t* id)
{
    if (!s || !req || !id)
        return TMC_E_INVALID;
    int rc = tmc::check_params(&req->params);
    if (rc != TMC_OK)
        return rc;
    if (req->params.shells > s->opt.max_shells)
        return TMC_E_INVALID;
    if (!tmc::photon_range_ok(req->first_photon, req->params.photons))
        return TMC_E_RANGE;
    const uint64_t words = 4 * (uint64_t)req->params.shells;
    if (words > s->opt.tally_words)
        return TMC_E_NOMEM;
    Request* r = new (std::nothrow) Request;
    if (!r)
        return TMC_E_NOMEM;
    r->set = *req;
    r->words = words;
    r->L = tmc::expected_interactions(&req->params, 0);
    r->chunk = chunk_photons(s->opt, r->L, std::max<uint64_t>(1, req->params.photons));
    std::lock_guard<std::mutex> lk(s->m);
    if (s->closing || s->error != TMC_OK) {
        delete r;
        return s->error != TMC_OK ? s->error : TMC_E_INVALID;
    }
    r->id = s->next_id++;
    r->t_submit = s->now();
    s->st.submitted++;
    s->state[r->id] = 0;
    if (req->params.photons == 0) {
        /* nothing to walk: finished at once with zero tallies */
        r->result.assign(words, 0);
        r->t_admit = r->t_done = r->t_submit;
        s->finished[r->id] = r;
        s->state[r->id] = 1;
        s->st.admitted++;
        s->st.completed++;
    } else {
        s->pending.push_back(r);
    }
    *id = r->id;
    s->cv_work.notify_one();
    return TMC_OK;
}

namespace {

int take_result(tmc_server* s, uint64_t id, uint64_t* tally, tmc_request_stats* st)
{
    auto it = s->finished.find(id);
    Request* r = it->second;
    if (tally)
        std::memcpy(tally, r->result.data(), r->words * sizeof(uint64_t));
    if (st) {
        st->photons = r->set.params.photons;
        st->interactions = r->interactions;
        st->submit_s = r->t_submit;
        st->admit_s = r->t_admit;
        st->done_s = r->t_done;
        st->latency_ms = (r->t_done - r->t_submit) * 1e3;
    }
    s->finished.erase(it);
    s->state.erase(id);
    delete r;
    return TMC_OK;
}

} // namespace

TMC_EXPORT int tmc_serve_poll(tmc_server* s, uint64_t id, uint64_t* tally, tmc_request_stats* st)
{
    if (!s)
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->state.find(id);
    if (it == s->state.end())
        return TMC_E_INVALID;
    if (it->second == 0)
        return s->error != TMC_OK ? s->error : TMC_PENDING;
    return take_result(s, id, tally, st);
}

TMC_EXPORT int tmc_serve_wait(tmc_server* s, uint64_t id, double timeout_s, uint64_t* tally, tmc_request_stats* st)
{
    if (!s)
        return TMC_E_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                             std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s));
    for (;;) {
        auto it = s->state.find(id);
        if (it == s->state.end())
            return TMC_E_INVALID;
        if (it->second == 1)
            return take_result(s, id, tally, st);
        if (s->error != TMC_OK)
            return s->error;
        if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && Clock::now() >= deadline)
            return TMC_PENDING;
    }
}

TMC_EXPORT int tmc_serve_get_stats(tmc_server* s, tmc_serve_stats* out)
{
    if (!s || !out)
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->st;
    out->chunks_claimed = s->claimed;
    out->pending = s->pending.size();
    out->active = s->active.size();
    out->elapsed_s = s->now();
    return s->error;
}

TMC_EXPORT int tmc_serve_close(tmc_server* s)
{
    if (!s)
        return TMC_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->closing = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable())
        s->worker.join();
    const int rc = s->error;
    server_free(s);
    return rc;
}
