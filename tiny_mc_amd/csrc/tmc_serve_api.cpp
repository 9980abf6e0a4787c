/* Serving: continuous batching of simulation requests on a resident
 * persistent kernel (include/tiny_mc_gpu.h, tmc_serve_*).
 *
 * Memory the server keeps for its lifetime:
 *   device      tally arena (128-bit sums of every resident request),
 *               per-slot done counters (photons, absorptions), claim cursor;
 *   device      also the request table (one BatchSet per slot), the chunk
 *               ring and the control words (published tail, stop flag),
 *               written by the server with host-to-device copies on its copy
 *               stream (descriptors before the tail that publishes them) and
 *               polled by the kernel with system-scope loads, so the polling
 *               never crosses the host bus;
 *   host        pinned staging mirrors of those tables and copy buffers.
 *
 * The server thread loops:
 *   1. admit queued requests in arrival order while a slot, room in the
 *      tally arena and room in the ring are free (zeroing the request's
 *      tally and counters with copies on the copy stream);
 *   2. publish chunks round-robin over the admitted requests, keeping a
 *      bounded number of published-but-unclaimed chunks, so a request that
 *      arrives while others run gets its first chunks claimed within about
 *      one chunk time of a wave (this is what folds it into the running
 *      kernel);
 *   3. (re)launch the kernel when published work is waiting and no kernel
 *      is running (the kernel leaves after idle_ms without work);
 *   4. copy the done counters back, and for every request whose photons are
 *      all counted copy its tally back, record its latency, free its slot.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
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

} // namespace

struct tmc_server {
    tmc_serve_options opt;
    int device = 0;
    uint32_t grid = 0, block = 0, total_waves = 0;
    hipStream_t kstream = nullptr, cstream = nullptr;
    /* device */
    unsigned long long* d_tally = nullptr;
    unsigned long long* d_done = nullptr;
    unsigned long long* d_counters = nullptr;
    unsigned long long* d_lost = nullptr; /* per wave: ticket held at exit */
    tmc::ServeControl* d_ctl = nullptr;
    tmc::BatchSet* d_sets = nullptr;
    tmc::BatchChunk* d_ring = nullptr;
    /* host, pinned staging mirrors of the device tables */
    uint64_t* h_ctl = nullptr; /* ring of staged control words */
    uint32_t ctl_next = 0;
    tmc::BatchSet* h_sets = nullptr;
    tmc::BatchChunk* h_ring = nullptr;
    /* host, pinned (copy staging) */
    unsigned long long* h_done = nullptr;
    unsigned long long* h_zero = nullptr;
    unsigned long long* h_stage = nullptr;
    unsigned long long* h_counters = nullptr;
    unsigned long long* h_lost = nullptr;
    size_t zero_words = 0;
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
    size_t rr = 0;
    uint64_t published = 0, claimed = 0;
    bool kernel_running = false;
    tmc_serve_stats st;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void server_free(tmc_server* s)
{
    if (s->kstream)
        (void)hipStreamSynchronize(s->kstream);
    if (s->cstream)
        (void)hipStreamSynchronize(s->cstream);
    if (s->d_tally)
        (void)hipFree(s->d_tally);
    if (s->d_done)
        (void)hipFree(s->d_done);
    if (s->d_counters)
        (void)hipFree(s->d_counters);
    if (s->d_lost)
        (void)hipFree(s->d_lost);
    for (void* d : { (void*)s->d_ctl, (void*)s->d_sets, (void*)s->d_ring })
        if (d)
            (void)hipFree(d);
    for (void* h : { (void*)s->h_ctl, (void*)s->h_sets, (void*)s->h_ring, (void*)s->h_done, (void*)s->h_zero,
                     (void*)s->h_stage, (void*)s->h_counters, (void*)s->h_lost })
        if (h)
            (void)hipHostFree(h);
    if (s->kstream)
        (void)hipStreamDestroy(s->kstream);
    if (s->cstream)
        (void)hipStreamDestroy(s->cstream);
    for (Request* r : s->pending)
        delete r;
    for (Request* r : s->active)
        delete r;
    for (auto& kv : s->finished)
        delete kv.second;
    delete s;
}

/* Chunk size of a request: about chunk_work absorptions, between 1 and
 * max_photons_per_lane photons per lane, never more than the request. */
uint64_t chunk_photons(const tmc_serve_options& o, double L, uint64_t N)
{
    double per_lane = std::ceil((double)o.chunk_work / (64.0 * L));
    per_lane = std::min(std::max(per_lane, 1.0), (double)o.max_photons_per_lane);
    return std::max<uint64_t>(1, std::min<uint64_t>(N, 64u * (uint64_t)per_lane));
}

constexpr uint32_t kCtlStage = 1024;

/* Copies a control word (tail or stop) to the device through its own
 * staging word, so a later update never changes a copy already queued. */
void put_ctl(tmc_server* s, size_t offset_words, uint64_t v)
{
    uint64_t* w = s->h_ctl + (s->ctl_next++ % kCtlStage);
    *w = v;
    if (hipMemcpyAsync((uint64_t*)s->d_ctl + offset_words, w, sizeof(uint64_t), hipMemcpyHostToDevice, s->cstream)
        != hipSuccess)
        s->error = TMC_E_HIP;
}

/* Copies ring entries [from, to) from the staging mirror to the device
 * ring, then the new tail that publishes them (same stream: in order). */
void publish(tmc_server* s, uint64_t from, uint64_t to)
{
    const uint64_t ring = s->opt.ring_chunks;
    uint64_t p = from;
    while (p < to) {
        const uint64_t i = p & (ring - 1);
        const uint64_t n = std::min(to - p, ring - i);
        if (hipMemcpyAsync(s->d_ring + i, s->h_ring + i, n * sizeof(tmc::BatchChunk), hipMemcpyHostToDevice,
                           s->cstream)
            != hipSuccess)
            s->error = TMC_E_HIP;
        p += n;
    }
    put_ctl(s, 0, to);
}

/* After a kernel has left: tickets [0, C) were handed out (C = claim
 * cursor); every one was walked except those a wave still held when it left
 * (d_lost).  Chunks published under a lost ticket are published again at the
 * tail; ring indices handed out beyond the tail are skipped. */
void recover_tickets(tmc_server* s)
{
    const uint64_t ring = s->opt.ring_chunks;
    if (hipMemcpyAsync(s->h_lost, s->d_lost, s->total_waves * sizeof(uint64_t), hipMemcpyDeviceToHost, s->cstream)
            != hipSuccess
        || hipMemcpyAsync(s->h_counters, s->d_counters, sizeof(uint64_t), hipMemcpyDeviceToHost, s->cstream)
            != hipSuccess
        || hipStreamSynchronize(s->cstream) != hipSuccess) {
        s->error = TMC_E_HIP;
        return;
    }
    const uint64_t C = s->h_counters[0];
    const uint64_t P0 = s->published;
    uint64_t pub = std::max(P0, C);
    for (uint32_t w = 0; w < s->total_waves; ++w) {
        const uint64_t t = s->h_lost[w];
        if (t == ~0ull || t >= P0)
            continue;
        s->h_ring[pub & (ring - 1)] = s->h_ring[t & (ring - 1)];
        ++pub;
    }
    s->claimed = C;
    if (pub != s->published) {
        const uint64_t from = std::max(P0, C);
        s->published = pub;
        publish(s, from, pub); /* indices handed out beyond the old tail are skipped */
    }
}

/* One pass of the server loop; called with s->m held. */
void serve_step(tmc_server* s)
{
    const uint64_t ring = s->opt.ring_chunks;
    /* 1. admit in arrival order */
    std::vector<Request*> admitted;
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
        tmc::BatchSet& d = s->h_sets[r->slot]; /* staging; copied below */
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
        std::memcpy(&d, &b, sizeof b);
        (void)hipMemcpyAsync(s->d_sets + r->slot, &d, sizeof d, hipMemcpyHostToDevice, s->cstream);
        (void)hipMemcpyAsync(s->d_tally + off, s->h_zero, r->words * sizeof(uint64_t), hipMemcpyHostToDevice,
                             s->cstream);
        (void)hipMemcpyAsync(s->d_done + 2 * (uint64_t)r->slot, s->h_zero, 2 * sizeof(uint64_t),
                             hipMemcpyHostToDevice, s->cstream);
        admitted.push_back(r);
        s->st.admitted++;
    }
    /* the copies above precede, on the copy stream, the tail that publishes
     * the new requests' first chunks */
    for (Request* r : admitted)
        s->active.push_back(r);

    /* 2. publish round-robin, bounded backlog */
    const uint64_t backlog_max = std::min<uint64_t>(ring - s->total_waves, 4ull * s->total_waves);
    bool any = true;
    const uint64_t before = s->published;
    while (any && s->published - s->claimed < backlog_max && !s->active.empty()) {
        any = false;
        for (size_t n = 0; n < s->active.size() && s->published - s->claimed < backlog_max; ++n) {
            Request* r = s->active[(s->rr + n) % s->active.size()];
            const uint64_t N = r->set.params.photons;
            if (r->next >= N)
                continue;
            const uint64_t cnt = std::min(r->chunk, N - r->next);
            tmc::BatchChunk& c = s->h_ring[s->published & (ring - 1)];
            c.set = r->slot | (r->gen << TMC_SERVE_SLOT_BITS);
            c.count = (uint32_t)cnt;
            c.first = r->next;
            r->next += cnt;
            s->published++;
            any = true;
        }
        s->rr = s->active.empty() ? 0 : (s->rr + 1) % s->active.size();
    }
    if (s->published != before) {
        publish(s, before, s->published);
        s->st.chunks_published = s->published;
    }

    /* 3. keep a kernel running while published work waits */
    if (s->kernel_running && hipStreamQuery(s->kstream) == hipSuccess) {
        s->kernel_running = false;
        recover_tickets(s);
    }
    if (!s->kernel_running && s->published > s->claimed && s->error == TMC_OK) {
        if (hipMemsetAsync(s->d_lost, 0xFF, s->total_waves * sizeof(uint64_t), s->kstream) != hipSuccess
            || tmc::launch_serve(s->opt.sampler, s->args, s->grid, s->block, s->kstream) != hipSuccess) {
            s->error = TMC_E_HIP;
        } else {
            s->kernel_running = true;
            s->st.kernel_launches++;
        }
    }

    /* 4. completions */
    if (s->active.empty())
        return;
    uint32_t slots_used = 0;
    for (Request* r : s->active)
        slots_used = std::max(slots_used, r->slot + 1);
    if (hipMemcpyAsync(s->h_done, s->d_done, 2 * (size_t)slots_used * sizeof(uint64_t), hipMemcpyDeviceToHost,
                       s->cstream) != hipSuccess
        || hipMemcpyAsync(s->h_counters, s->d_counters, sizeof(uint64_t), hipMemcpyDeviceToHost, s->cstream)
            != hipSuccess
        || hipStreamSynchronize(s->cstream) != hipSuccess) {
        s->error = TMC_E_HIP;
        return;
    }
    s->claimed = std::max<uint64_t>(s->claimed, s->h_counters[0]);
    std::vector<Request*> done;
    for (Request* r : s->active)
        if (r->next >= r->set.params.photons && s->h_done[2 * (uint64_t)r->slot] >= r->set.params.photons)
            done.push_back(r);
    if (done.empty())
        return;
    uint64_t stage = 0;
    for (Request* r : done) {
        (void)hipMemcpyAsync(s->h_stage + stage, s->d_tally + r->tally_off, r->words * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, s->cstream);
        stage += r->words;
    }
    if (hipStreamSynchronize(s->cstream) != hipSuccess) {
        s->error = TMC_E_HIP;
        return;
    }
    stage = 0;
    const double t = s->now();
    for (Request* r : done) {
        r->result.assign(s->h_stage + stage, s->h_stage + stage + r->words);
        stage += r->words;
        r->interactions = s->h_done[2 * (uint64_t)r->slot + 1];
        r->t_done = t;
        s->arena.release(r->tally_off, r->words);
        s->free_slots.push_back(r->slot);
        s->active.erase(std::find(s->active.begin(), s->active.end(), r));
        s->finished[r->id] = r;
        s->state[r->id] = 1;
        s->st.completed++;
        s->st.photons_completed += r->set.params.photons;
        s->st.interactions_completed += r->interactions;
    }
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
            if (s->kernel_running && hipStreamQuery(s->kstream) == hipSuccess) {
                s->kernel_running = false;
                recover_tickets(s);
            }
            continue;
        }
        serve_step(s);
        if (s->error != TMC_OK) {
            s->cv_done.notify_all();
            break;
        }
        /* short pause: other threads submit and poll under the lock */
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(100));
        lk.lock();
    }
    put_ctl(s, 1, 1);
    (void)hipStreamSynchronize(s->cstream);
    lk.unlock();
    (void)hipStreamSynchronize(s->kstream);
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
    if (o.sampler > TMC_SAMPLER_REJECTION)
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
        o.chunk_work = 1ull << 20;
    if (!o.max_photons_per_lane)
        o.max_photons_per_lane = 16;
    if (!o.idle_ms)
        o.idle_ms = 50;
    if (!o.tally_words)
        o.tally_words = (uint64_t)o.max_requests * 4 * o.max_shells;
    if (o.max_shells > TMC_MAX_SHELLS || o.max_requests > (1u << TMC_SERVE_SLOT_BITS) || o.block % 64
        || o.block > 1024 || (o.ring_chunks & (o.ring_chunks - 1)) || o.ring_chunks > (1ull << 31))
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
        || tmc::serve_occupancy(o.sampler, o.block, o.max_shells, 1, &wg) != hipSuccess || wg <= 0) {
        delete s;
        return TMC_E_INVALID;
    }
    s->block = o.block;
    s->grid = (uint32_t)wg * (uint32_t)c->compute_units;
    s->total_waves = s->grid * (o.block / 64);
    if (o.ring_chunks <= 2ull * s->total_waves) {
        delete s;
        return TMC_E_INVALID;
    }
    s->zero_words = std::max<size_t>(4 * (size_t)o.max_shells, 2);
    const size_t R = o.max_requests;
    if (hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking) != hipSuccess
        || hipStreamCreateWithFlags(&s->cstream, hipStreamNonBlocking) != hipSuccess
        || hipMalloc((void**)&s->d_tally, o.tally_words * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_done, 2 * R * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_counters, 4 * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&s->d_lost, s->total_waves * sizeof(uint64_t)) != hipSuccess
        || hipHostMalloc((void**)&s->h_lost, s->total_waves * sizeof(uint64_t), 0) != hipSuccess
        || hipMalloc((void**)&s->d_ctl, sizeof(tmc::ServeControl)) != hipSuccess
        || hipMalloc((void**)&s->d_sets, R * sizeof(tmc::BatchSet)) != hipSuccess
        || hipMalloc((void**)&s->d_ring, o.ring_chunks * sizeof(tmc::BatchChunk)) != hipSuccess
        || hipHostMalloc((void**)&s->h_ctl, kCtlStage * sizeof(uint64_t), 0) != hipSuccess
        || hipHostMalloc((void**)&s->h_sets, R * sizeof(tmc::BatchSet), 0) != hipSuccess
        || hipHostMalloc((void**)&s->h_ring, o.ring_chunks * sizeof(tmc::BatchChunk), 0) != hipSuccess
        || hipHostMalloc((void**)&s->h_done, 2 * R * sizeof(uint64_t), 0) != hipSuccess
        || hipHostMalloc((void**)&s->h_zero, s->zero_words * sizeof(uint64_t), 0) != hipSuccess
        || hipHostMalloc((void**)&s->h_stage, std::max<uint64_t>(o.tally_words, 1) * sizeof(uint64_t), 0)
            != hipSuccess
        || hipHostMalloc((void**)&s->h_counters, 4 * sizeof(uint64_t), 0) != hipSuccess) {
        server_free(s);
        return TMC_E_NOMEM;
    }
    std::memset(s->h_ctl, 0, kCtlStage * sizeof(uint64_t));
    std::memset(s->h_sets, 0, R * sizeof(tmc::BatchSet));
    std::memset(s->h_zero, 0, s->zero_words * sizeof(uint64_t));
    const unsigned long long zeros[4] = { 0, 0, 0, 0 };
    if (hipMemcpy(s->d_counters, zeros, sizeof zeros, hipMemcpyHostToDevice) != hipSuccess
        || hipMemset(s->d_done, 0, 2 * R * sizeof(uint64_t)) != hipSuccess
        || hipMemset(s->d_ctl, 0, sizeof(tmc::ServeControl)) != hipSuccess) {
        server_free(s);
        return TMC_E_HIP;
    }
    s->arena.init(o.tally_words);
    s->slot_gen.assign(R, 0);
    for (uint32_t i = 0; i < R; ++i)
        s->free_slots.push_back((uint32_t)(R - 1 - i));

    std::memset(&s->args, 0, sizeof s->args);
    s->args.sets = s->d_sets;
    s->args.chunks = s->d_ring;
    s->args.nchunks = 0;
    s->args.shells_max = o.max_shells;
    s->args.copies = 1;
    s->args.flush_interval = 1u << 14;
    s->args.ring_mask = (uint32_t)(o.ring_chunks - 1);
    s->args.counters = s->d_counters;
    s->args.tally = s->d_tally;
    s->args.done = s->d_done;
    s->args.ctl = s->d_ctl;
    s->args.idle_ticks = (uint64_t)o.idle_ms * 100000ull; /* 100 MHz realtime counter */
    s->args.lost = s->d_lost;

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
    if (req->params.photons > UINT64_MAX - req->first_photon)
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
