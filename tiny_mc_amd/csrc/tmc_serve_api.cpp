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
 * runs: it writes descriptors, ring entries and the tail with ordinary
 * stores (release where the kernel must see earlier stores first) and reads
 * finished requests from the mirror once their flag holds their token.  (A
 * copy could sit behind another process's or thread's kernel that in turn
 * waits for the compute units the resident kernel holds.)
 *
 * The server thread loops:
 *   1. notice a kernel that has left (it leaves only when asked to stop);
 *   2. cancel requests whose deadline has passed;
 *   3. admit queued requests in arrival order while a slot and room in the
 *      tally arena are free;
 *   4. publish ring entries: first the cancellations of admitted requests,
 *      then chunks, shortest remaining work first, keeping about one
 *      published-but-unclaimed chunk per resident wave, so a request that
 *      arrives while others run gets its first chunks claimed within about
 *      one chunk time of a wave (this is what folds it into the running
 *      kernel).  A chunk carries about chunk_work absorptions while the
 *      unpublished work is small; as it grows beyond a few chunks per wave,
 *      chunks grow with it (up to max_chunk_scale times), so a loaded server
 *      spends fewer wave drains per absorption.  One request counts towards
 *      that backlog with at most TMC_SERVE_REQUEST_SCALE times the
 *      threshold, so a large request alone does not take every wave for a
 *      long chunk;
 *   5. launch the kernel when published work waits and none is running;
 *   6. finish every request whose flag holds its token: take its sums from
 *      the mirror (or drop them, for a cancelled request), record its
 *      latency, free its slot and arena region.
 * After idle_ms with no request in the server it asks the kernel to stop:
 * then every published chunk has been claimed and done, so each wave's
 * ticket lies past the tail and the next launch starts its tickets at the
 * tail.
 *
 * Ring protocol (RingEntry): the host writes an entry's fields, then its
 * sequence number (release), and reuses a slot only after the wave that
 * read the previous entry has marked it consumed; the tail (release, after
 * the entries) tells waves where to look.
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
    uint64_t next = 0; /* next unpublished photon offset */
    uint64_t skip = 0; /* cancelled photons not yet announced to the kernel */
    double L = 1.0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double deadline = 0; /* seconds since the server opened; 0: none */
    int cancelled = 0; /* TMC_E_CANCELLED or TMC_E_DEADLINE once cancelled */
    uint64_t interactions = 0;
    std::vector<uint64_t> result;
};

/* Workgroup slots the serving kernel leaves free beyond reserve_per_cu per
 * compute unit: room for the runtime's one-workgroup kernels.  A kernel of
 * many workgroups needs more (measured on MI355X, 256 CUs, beside a resident
 * kernel holding 7 of its 7 workgroup slots per CU but 8 in all: a
 * 1-workgroup kernel ran at once, a 16-workgroup one waited for the
 * resident kernel to end; with 2 slots per CU free a 1024-workgroup one ran
 * at once), hence reserve_per_cu. */
constexpr uint32_t kFreeSlots = 8;
/* Minimum time between two reads of the host's control words by the
 * kernel, in 100 MHz realtime ticks (2 us). */
constexpr uint32_t kRefreshTicks = 200;
/* Chunks grow once the unpublished work exceeds this many base chunks per
 * resident wave. */
constexpr double kChunksPerWave = 2.0;
/* One request counts towards that backlog with at most this many times the
 * threshold.  Without a cap, a single large request arriving at a lightly
 * loaded server sizes its own chunks to half of its work per wave, every
 * wave takes one of them at about the same time, and the requests arriving
 * behind it wait for the first of those chunks to end.  Measured on MI355X
 * (request mix of Pareto(1.5) photon counts, 0.5 load; p99 on two request
 * lists, and the rate at saturation): no cap 148 / 130 ms, 0.943; 8x 125 /
 * 90 ms, 0.932; 4x 95 / 74 ms, 0.933; 2x 78 / 72 ms, 0.916; 1x 78 / 71 ms,
 * 0.896. */
#ifndef TMC_SERVE_REQUEST_SCALE
#define TMC_SERVE_REQUEST_SCALE 2
#endif

/* Request states in tmc_server::state besides the negative statuses of
 * cancelled requests. */
constexpr int kQueuedOrActive = 0;
constexpr int kFinished = 1;

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
    tmc::RingEntry* h_ring = nullptr;
    unsigned long long* h_out = nullptr;
    unsigned long long* h_flags = nullptr;
    tmc::BatchArgs args;

    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool closing = false;
    double close_deadline = -1; /* tmc_serve_close_timeout: cancel what is unfinished at this time */
    bool worker_done = false; /* the worker has stopped the kernel and left its loop */
    int error = TMC_OK;
    Clock::time_point t0;

    uint64_t next_id = 1;
    std::deque<Request*> pending;
    std::vector<Request*> active; /* admitted, unfinished */
    std::map<uint64_t, Request*> finished;
    std::map<uint64_t, int> state; /* id -> kQueuedOrActive, kFinished or the status of a cancelled request */
    std::vector<uint32_t> free_slots;
    std::vector<uint32_t> slot_gen;
    tmc::Arena arena;
    double unpublished_work = 0; /* expected absorptions of the photons not yet published,
                                    each request's at most request_work_cap */
    double request_work_cap = 0;
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

/* Frees a server whose kernel is not running.  Its device and pinned host
 * memory and its stream go through tmc::release, which defers them while
 * another server's kernel is resident (hipHostFree would wait for that
 * kernel), and which first waits for the server's stream: the clears queued
 * at open may still sit behind that kernel, on a hardware queue HIP shares
 * between streams. */
void server_free(tmc_server* s)
{
    void* dev[] = { s->d_tally, s->d_done, s->d_counters, s->d_cache };
    void* host[] = { s->h_ctl, s->h_sets, s->h_ring, s->h_out, s->h_flags };
    hipStream_t k = s->kstream;
    tmc::release([k, dev, host]() {
        if (k)
            (void)hipStreamSynchronize(k);
        for (void* d : dev)
            if (d)
                (void)hipFree(d);
        for (void* h : host)
            if (h)
                (void)hipHostFree(h);
        if (k)
            (void)hipStreamDestroy(k);
    });
    for (Request* r : s->pending)
        delete r;
    for (Request* r : s->active)
        delete r;
    for (auto& kv : s->finished)
        delete kv.second;
    delete s;
}

void store_release(uint64_t* p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
uint64_t load_acquire(const uint64_t* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }

/* What `photons` unpublished photons of a request with L expected
 * absorptions each add to the backlog that sizes chunks. */
double backlog_share(const tmc_server* s, uint64_t photons, double L)
{
    return tmc::backlog_share(photons, L, s->request_work_cap);
}

/* Writes ring entry s->published if its slot is free: the entry `ring`
 * indices before it (if any) has been read by its wave.  False if not (the
 * caller publishes later). */
bool publish_entry(tmc_server* s, uint32_t set, uint32_t count, uint64_t first)
{
    const uint64_t ring = s->opt.ring_chunks;
    const uint64_t t = s->published;
    tmc::RingEntry& e = s->h_ring[t & (ring - 1)];
    if (t >= ring && load_acquire(&e.seq) != ((t - ring) | TMC_RING_CONSUMED))
        return false;
    e.set = set;
    e.count = count;
    e.first = first;
    store_release(&e.seq, t);
    s->published++;
    return true;
}

/* Asks the running kernel to stop; called only when no request is admitted
 * or queued and nothing will be published until it has left. */
void request_stop(tmc_server* s)
{
    if (s->kernel_running && !s->stopping) {
        store_release(&s->h_ctl->stop, s->launch_id);
        s->stopping = true;
    }
}

/* The kernel has ended (stream idle): next launch starts at the tail. */
void kernel_left(tmc_server* s, hipError_t q)
{
    s->kernel_running = false;
    s->stopping = false;
    tmc::serving_kernel_left();
    if (q != hipSuccess) {
        s->error = TMC_E_HIP;
        return;
    }
    s->base = s->published;
    s->claimed = std::max(s->claimed, s->published);
}

/* Notices a kernel that has left.  It leaves only after a stop request,
 * when every published chunk had been claimed. */
void reap_kernel(tmc_server* s)
{
    if (!s->kernel_running)
        return;
    const hipError_t q = hipStreamQuery(s->kstream);
    if (q == hipErrorNotReady)
        return;
    kernel_left(s, q);
}

/* Cancels request r (any state but handed over) with status `why`
 * (TMC_E_CANCELLED or TMC_E_DEADLINE).  A queued request is dropped at
 * once.  An admitted one stops publishing: its unpublished photons go to
 * the kernel as one cancellation entry, so the request finishes, and frees
 * its slot and arena region, as soon as the chunks already published are
 * done; its sums are then dropped. */
void cancel_request(tmc_server* s, Request* r, int why, bool queued)
{
    s->state[r->id] = why;
    s->cv_done.notify_all(); /* a waiting caller gets the status now */
    if (why == TMC_E_DEADLINE)
        s->st.expired++;
    else
        s->st.cancelled++;
    if (queued) {
        s->unpublished_work -= backlog_share(s, r->set.params.photons, r->L);
        auto it = std::find(s->pending.begin(), s->pending.end(), r);
        if (it != s->pending.end())
            s->pending.erase(it);
        delete r;
        return;
    }
    r->cancelled = why;
    const uint64_t N = r->set.params.photons;
    if (r->next < N) {
        s->unpublished_work -= backlog_share(s, N - r->next, r->L);
        r->skip = N - r->next;
        r->next = N;
    }
}

/* One pass of the server loop; called with s->m held. */
void serve_step(tmc_server* s)
{
    reap_kernel(s);
    if (s->error != TMC_OK)
        return;
    const double tnow = s->now();

    /* deadlines; and at a bounded close's deadline, everything */
    const bool close_now = s->closing && s->close_deadline >= 0 && tnow >= s->close_deadline;
    for (size_t i = 0; i < s->pending.size();) {
        Request* r = s->pending[i];
        if (close_now || (r->deadline > 0 && tnow >= r->deadline))
            cancel_request(s, r, close_now ? TMC_E_CANCELLED : TMC_E_DEADLINE, true);
        else
            ++i;
    }
    for (Request* r : s->active)
        if (!r->cancelled && (close_now || (r->deadline > 0 && tnow >= r->deadline)))
            cancel_request(s, r, close_now ? TMC_E_CANCELLED : TMC_E_DEADLINE, false);

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
        r->t_admit = tnow;
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
        /* read by the kernel only after an entry of this request is
         * published, which the release store of its sequence number orders
         * after */
        std::memcpy(&s->h_sets[r->slot], &b, sizeof b);
        s->active.push_back(r);
        s->st.admitted++;
    }

    /* publish: cancellations first, then chunks, shortest remaining work
     * first, keeping about one unclaimed chunk per resident wave.  A slot of
     * the ring is rewritten only after its previous entry was read
     * (publish_entry). */
    s->claimed = std::max<uint64_t>(s->claimed, load_acquire(&s->h_ctl->claimed));
    const uint64_t backlog_max = std::min<uint64_t>(s->opt.ring_chunks, s->total_waves);
    auto unclaimed = [s]() -> uint64_t { return s->published > s->claimed ? s->published - s->claimed : 0; };
    const uint64_t before = s->published;
    if (!(s->kernel_running && s->stopping)) {
        for (Request* r : s->active) {
            if (r->skip && publish_entry(s, r->slot | (r->gen << TMC_SERVE_SLOT_BITS), 0u, r->skip))
                r->skip = 0;
        }
        /* chunk size: the base work while the unpublished work is within
         * kChunksPerWave base chunks per wave, growing with it beyond (one
         * request alone to at most TMC_SERVE_REQUEST_SCALE times) */
        const double base_work = (double)s->opt.chunk_work;
        const double scale = tmc::serve_chunk_scale(s->unpublished_work, base_work, s->total_waves, kChunksPerWave,
                                                    s->opt.max_chunk_scale);
        const double work = base_work * scale;
        const uint64_t max_per_lane = (uint64_t)std::min<double>(
            (double)(1u << 24), std::floor((double)s->opt.max_photons_per_lane * scale));
        std::vector<Request*> order;
        if (unclaimed() < backlog_max) {
            for (Request* r : s->active)
                if (r->next < r->set.params.photons)
                    order.push_back(r);
            std::sort(order.begin(), order.end(), [](const Request* x, const Request* y) {
                return (double)(x->set.params.photons - x->next) * x->L
                    < (double)(y->set.params.photons - y->next) * y->L;
            });
        }
        bool ring_full = false;
        for (Request* r : order) {
            const uint64_t N = r->set.params.photons;
            while (r->next < N && unclaimed() < backlog_max) {
                const uint64_t cnt = tmc::chunk_photons(work, r->L, max_per_lane, N - r->next);
                if (!publish_entry(s, r->slot | (r->gen << TMC_SERVE_SLOT_BITS), (uint32_t)cnt, r->next)) {
                    ring_full = true;
                    break;
                }
                s->unpublished_work -= backlog_share(s, N - r->next, r->L) - backlog_share(s, N - r->next - cnt, r->L);
                r->next += cnt;
            }
            if (ring_full || unclaimed() >= backlog_max)
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
        tmc::serving_kernel_started();
        if (tmc::launch_serve(s->opt.sampler, s->opt.math, s->args, s->grid, s->block, s->kstream) != hipSuccess) {
            tmc::serving_kernel_left();
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
        if (r->next < r->set.params.photons || r->skip
            || load_acquire((const uint64_t*)s->h_flags + 2 * r->slot) != r->id) {
            ++i;
            continue;
        }
        s->arena.release(r->tally_off, r->words);
        s->free_slots.push_back(r->slot);
        s->active[i] = s->active.back();
        s->active.pop_back();
        if (r->cancelled) {
            delete r; /* its state already holds the status to report */
            continue;
        }
        const unsigned long long* src = s->h_out + r->tally_off;
        r->result.assign(src, src + r->words);
        r->interactions = s->h_flags[2 * (uint64_t)r->slot + 1];
        r->t_done = t;
        s->finished[r->id] = r;
        s->state[r->id] = kFinished;
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
    /* every request is finished or cancelled (or the server failed): let
     * the kernel go */
    if (s->kernel_running) {
        store_release(&s->h_ctl->stop, s->launch_id);
        hipError_t q;
        for (;;) {
            q = hipStreamQuery(s->kstream);
            if (q != hipErrorNotReady)
                break;
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(200));
            lk.lock();
        }
        kernel_left(s, q);
    }
    s->worker_done = true;
    s->cv_done.notify_all();
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
    /* 128: at saturation (Poisson arrivals of the sweep's request mix at the
     * batch kernel's rate) the served rate measured 0.76 of the batch rate
     * with chunks up to 32x, 0.81 with 128x and 512x; light-load latency is
     * the same (the scale stays 1 while the backlog is small) */
    if (!o.max_chunk_scale)
        o.max_chunk_scale = 128;
    if (!o.tally_words)
        o.tally_words = (uint64_t)o.max_requests * 4 * o.max_shells;
    /* a chunk (at most 64 * 2^24 photons) keeps its photon count in 32 bits */
    if (o.max_shells > TMC_MAX_SHELLS || o.max_requests > (1u << TMC_SERVE_SLOT_BITS) || o.block % 64
        || o.block > 1024 || (o.ring_chunks & (o.ring_chunks - 1)) || o.ring_chunks > (1ull << 31)
        || o.ring_chunks < 2 || o.max_photons_per_lane > (1u << 24) || o.max_chunk_scale > 1024 || o.reserved)
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
    if ((uint32_t)wg <= o.reserve_per_cu) {
        delete s;
        return TMC_E_INVALID; /* nothing left to serve with */
    }
    s->block = o.block;
    s->grid = ((uint32_t)wg - o.reserve_per_cu) * (uint32_t)c->compute_units;
    s->grid = s->grid > 2 * kFreeSlots ? s->grid - kFreeSlots : (s->grid + 1) / 2;
    s->total_waves = s->grid * (o.block / 64);
    s->request_work_cap = TMC_SERVE_REQUEST_SCALE * kChunksPerWave * (double)o.chunk_work * (double)s->total_waves;
    const size_t R = o.max_requests;
    if (hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking) != hipSuccess
        || hipMalloc((void**)&s->d_tally, o.tally_words * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_done, 2 * R * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_counters, 4 * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_cache, (TMC_SERVE_CACHE_COPIES + 1) * sizeof(tmc::ServeCache)) != hipSuccess
        || !host_alloc(&s->h_ctl, 1) || !host_alloc(&s->h_sets, R) || !host_alloc(&s->h_ring, o.ring_chunks)
        || !host_alloc(&s->h_out, o.tally_words) || !host_alloc(&s->h_flags, 2 * R)) {
        server_free(s);
        return TMC_E_NOMEM;
    }
    /* no ticket equals an unwritten entry's sequence number */
    for (uint64_t i = 0; i < o.ring_chunks; ++i)
        s->h_ring[i].seq = ~0ull;
    /* queued on the server's own stream, ahead of its first kernel, and
     * not waited for: HIP maps streams onto a few hardware queues, and this
     * stream may share one with another server's resident kernel */
    if (hipMemsetAsync(s->d_tally, 0, o.tally_words * sizeof(uint64_t), s->kstream) != hipSuccess
        || hipMemsetAsync(s->d_done, 0, 2 * R * sizeof(uint64_t), s->kstream) != hipSuccess
        || hipMemsetAsync(s->d_counters, 0, 4 * sizeof(uint64_t), s->kstream) != hipSuccess
        || hipMemsetAsync(s->d_cache, 0, (TMC_SERVE_CACHE_COPIES + 1) * sizeof(tmc::ServeCache), s->kstream)
            != hipSuccess) {
        server_free(s);
        return TMC_E_HIP;
    }
    s->arena.init(o.tally_words);
    s->slot_gen.assign(R, 0);
    for (uint32_t i = 0; i < R; ++i)
        s->free_slots.push_back((uint32_t)(R - 1 - i));

    std::memset(&s->args, 0, sizeof s->args);
    s->args.sets = device_view(s->h_sets);
    s->args.ring = device_view(s->h_ring);
    s->args.hctl = device_view(s->h_ctl);
    s->args.out = device_view(s->h_out);
    s->args.flags = device_view(s->h_flags);
    if (!s->args.sets || !s->args.ring || !s->args.hctl || !s->args.out || !s->args.flags) {
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

TMC_EXPORT int tmc_serve_submit_ex(tmc_server* s, const tmc_set* req, double deadline_s, uint64_t* id)
{
    if (!s || !req || !id || std::isnan(deadline_s))
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
    std::lock_guard<std::mutex> lk(s->m);
    if (s->closing || s->error != TMC_OK) {
        delete r;
        return s->error != TMC_OK ? s->error : TMC_E_INVALID;
    }
    r->id = s->next_id++;
    r->t_submit = s->now();
    if (deadline_s > 0)
        r->deadline = r->t_submit + deadline_s;
    s->st.submitted++;
    s->state[r->id] = kQueuedOrActive;
    if (req->params.photons == 0) {
        /* nothing to walk: finished at once with zero tallies */
        r->result.assign(words, 0);
        r->t_admit = r->t_done = r->t_submit;
        s->finished[r->id] = r;
        s->state[r->id] = kFinished;
        s->st.admitted++;
        s->st.completed++;
    } else {
        s->unpublished_work += backlog_share(s, req->params.photons, r->L);
        s->pending.push_back(r);
    }
    *id = r->id;
    s->cv_work.notify_one();
    return TMC_OK;
}

TMC_EXPORT int tmc_serve_submit(tmc_server* s, const tmc_set* req, uint64_t* id)
{
    return tmc_serve_submit_ex(s, req, 0.0, id);
}

TMC_EXPORT int tmc_serve_cancel(tmc_server* s, uint64_t id)
{
    if (!s)
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->state.find(id);
    if (it == s->state.end())
        return TMC_E_INVALID;
    if (it->second < 0)
        return TMC_OK; /* cancelled already */
    if (it->second == kFinished) {
        auto f = s->finished.find(id);
        delete f->second;
        s->finished.erase(f);
        it->second = TMC_E_CANCELLED;
        s->st.cancelled++;
        return TMC_OK;
    }
    for (Request* r : s->pending)
        if (r->id == id) {
            cancel_request(s, r, TMC_E_CANCELLED, true);
            return TMC_OK;
        }
    for (Request* r : s->active)
        if (r->id == id) {
            cancel_request(s, r, TMC_E_CANCELLED, false);
            s->cv_work.notify_one();
            return TMC_OK;
        }
    return TMC_E_INVALID;
}

namespace {

/* Hands a request's outcome over (once): its tally and statistics, or the
 * status of its cancellation. */
int take_result(tmc_server* s, std::map<uint64_t, int>::iterator st_it, uint64_t* tally, tmc_request_stats* st)
{
    const uint64_t id = st_it->first;
    if (st_it->second < 0) {
        const int why = st_it->second;
        s->state.erase(st_it);
        return why;
    }
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
    s->state.erase(st_it);
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
    if (it->second == kQueuedOrActive)
        return s->error != TMC_OK ? s->error : TMC_PENDING;
    return take_result(s, it, tally, st);
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
        if (it->second != kQueuedOrActive)
            return take_result(s, it, tally, st);
        if (s->error != TMC_OK)
            return s->error;
        if (Clock::now() >= deadline)
            return TMC_PENDING;
        s->cv_done.wait_until(lk, deadline);
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
    out->grid = s->grid;
    out->block = s->block;
    return s->error;
}

TMC_EXPORT int tmc_serve_close_timeout(tmc_server* s, double timeout_s)
{
    if (!s || std::isnan(timeout_s))
        return TMC_E_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    const auto start = Clock::now();
    if (!s->closing || (timeout_s >= 0 && s->close_deadline < 0)) {
        s->closing = true;
        if (timeout_s >= 0)
            s->close_deadline = s->now() + timeout_s;
        s->cv_work.notify_all();
    }
    /* the requests get timeout_s to finish, the cancelled ones as long
     * again (and one second) to drain and the kernel to leave */
    const auto until = start + std::chrono::duration_cast<Clock::duration>(
                                   std::chrono::duration<double>(timeout_s < 0 ? 1e9 : 2 * timeout_s + 1.0));
    while (!s->worker_done) {
        if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout && Clock::now() >= until && !s->worker_done)
            return TMC_E_TIMEOUT; /* the server stays open: call again */
    }
    lk.unlock();
    if (s->worker.joinable())
        s->worker.join();
    const int rc = s->error;
    server_free(s);
    return rc;
}

TMC_EXPORT int tmc_serve_close(tmc_server* s)
{
    return tmc_serve_close_timeout(s, -1.0);
}
