/* Host side of the batched sweep (include/tiny_mc_gpu.h, tmc_batch_*).
 *
 * A plan owns the device copies of the set table and the chunk queue, the
 * tally pool and, optionally, a HIP graph of the steady launch sequence
 * (clear kernel, persistent walk kernel, copy-back), so repeated runs of the
 * same batch cost one graph launch.
 *
 * Scheduling: each set's expected absorptions per photon L follow from the
 * binary32 weight cadence (tmc::expected_interactions).  A set of N photons
 * is cut into chunks of 64 * P photons with P = ceil(chunk_work / (64 L)),
 * so every chunk carries about chunk_work absorptions and a wave drains (its
 * lanes run out of the set's photons) once per ~P photons per lane.  Chunks
 * are sorted by expected duration, longest first, which keeps the long
 * high-albedo chunks off the end of the launch.
 *
 * Passes: a wave's LDS slot holds the largest shell count of its launch, so
 * a few sets with many shells would lower every wave's occupancy.  Sets are
 * grouped by the launch shape their shell count allows (classes with under a
 * tenth of the work join the next larger class), and each group runs as a
 * launch of its own, one after the other on the same stream, with its own
 * chunk queue and cursor; the tallies and counters are shared.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/tiny_mc_gpu.h"
#include "tmc_batch_internal.h"
#include "tmc_context.h"

#define TMC_EXPORT extern "C" __attribute__((visibility("default")))

static_assert(TMC_BATCH_CURSOR0 + TMC_BATCH_MAX_PASSES <= TMC_BATCH_COUNTERS, "pass cursors beyond the counters");

namespace {

/* One launch of a batch: the sets of one LDS size class. */
struct PassPlan {
    uint32_t grid = 0, block = 0, wps = 0, shells_max = 1;
    uint64_t nchunks = 0;
    uint64_t chunk_work = 0;
    double work = 0.0;
    tmc::BatchArgs args;
};

} // namespace

struct tmc_batch_plan {
    uint32_t nsets = 0;
    uint64_t tally_words = 0;
    uint64_t nchunks = 0;
    uint64_t photons = 0;
    double expected = 0.0;
    /* shape of the pass with the most work (reported in the statistics) */
    uint64_t chunk_work = 0;
    uint32_t sampler = 0, math = 0, grid = 0, block = 0, copies = 1, wps = 0, shells_max = 1;
    uint32_t npasses = 0;
    PassPlan pass[TMC_BATCH_MAX_PASSES];
    bool use_graph = false;
    tmc::BatchSet* d_sets = nullptr;
    tmc::BatchChunk* d_chunks = nullptr;
    unsigned long long* d_tally = nullptr;
    unsigned long long* d_done = nullptr;
    unsigned long long* d_counters = nullptr;
    unsigned long long* h_tally = nullptr; /* pinned */
    unsigned long long* h_counters = nullptr; /* pinned, TMC_BATCH_COUNTERS words */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

/* Device and pinned memory, and the graph, go through tmc::release, which
 * defers them while a serving kernel is resident (hipFree and hipHostFree
 * would wait for it). */
void plan_free(tmc_batch_plan* p)
{
    if (!p)
        return;
    if (p->exec || p->graph) {
        hipGraphExec_t exec = p->exec;
        hipGraph_t graph = p->graph;
        tmc::release([exec, graph]() {
            if (exec)
                (void)hipGraphExecDestroy(exec);
            if (graph)
                (void)hipGraphDestroy(graph);
        });
    }
    for (void* d : { (void*)p->d_sets, (void*)p->d_chunks, (void*)p->d_tally, (void*)p->d_done, (void*)p->d_counters })
        tmc::release_device(d);
    tmc::release_host(p->h_tally);
    tmc::release_host(p->h_counters);
    delete p;
}

/* Enqueues clear + walk + copies on `stream`. */
hipError_t enqueue_run(tmc_batch_plan* p, hipStream_t stream)
{
    hipError_t e = tmc::launch_batch_clear(p->d_tally, p->tally_words, p->d_done, p->nsets, p->d_counters, stream);
    if (e != hipSuccess)
        return e;
    for (uint32_t j = 0; j < p->npasses; ++j) {
        const PassPlan& q = p->pass[j];
        if (!q.nchunks)
            continue;
        e = tmc::launch_batch(p->sampler, p->math, q.args, q.grid, q.block, stream);
        if (e != hipSuccess)
            return e;
    }
    e = hipMemcpyAsync(p->h_tally, p->d_tally, p->tally_words * sizeof(uint64_t), hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess)
        return e;
    return hipMemcpyAsync(p->h_counters, p->d_counters, TMC_BATCH_COUNTERS * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          stream);
}

/* Launch shape for per-wave LDS slots of shells_max shells: the requested
 * block, or the largest of 256/128/64 whose slots fit on a CU, with the
 * resident workgroups capped by waves_per_simd when requested.  *fits is
 * false when no block fits. */
hipError_t launch_shape(const tmc_batch_plan* p, const tmc_batch_options* o, uint32_t cus, uint32_t shells_max,
                        PassPlan* q, bool* fits)
{
    *fits = false;
    std::vector<uint32_t> cand;
    if (o->block)
        cand.push_back(o->block);
    else
        cand = { 256u, 128u, 64u };
    for (uint32_t b : cand) {
        int n = 0;
        if (tmc::batch_lds_bytes(b, shells_max, p->copies) <= 160u * 1024u) {
            const hipError_t e = tmc::batch_occupancy(p->sampler, p->math, b, shells_max, p->copies, &n);
            if (e != hipSuccess)
                return e;
        }
        if (n <= 0)
            continue;
        const uint32_t waves_per_wg = b / 64u;
        if (o->waves_per_simd)
            n = std::min<int>(n, (int)std::max(1u, o->waves_per_simd * 4u / waves_per_wg));
        q->block = b;
        q->grid = (uint32_t)n * cus;
        q->wps = (uint32_t)n * waves_per_wg / 4u;
        q->shells_max = shells_max;
        *fits = true;
        return hipSuccess;
    }
    return hipSuccess;
}

int plan_create_locked(tmc::Context* c, const tmc_set* sets, uint32_t nsets, const tmc_batch_options* o,
                       tmc_batch_plan** out)
{
    if (o->sampler > TMC_SAMPLER_REJECTION || o->math > TMC_MATH_OWN)
        return TMC_E_INVALID;
    if (o->block && (o->block < 64u || o->block > 1024u || o->block % 64u))
        return TMC_E_INVALID;
    const uint32_t copies = o->copies ? o->copies : 1u;
    if (copies != 1u && copies != 2u)
        return TMC_E_INVALID;
    if (o->waves_per_simd > 8u)
        return TMC_E_INVALID;

    tmc_batch_plan* p = new (std::nothrow) tmc_batch_plan;
    if (!p)
        return TMC_E_NOMEM;
    p->nsets = nsets;
    p->sampler = o->sampler;
    p->math = o->math;
    p->copies = copies;
    p->use_graph = o->use_graph != 0;

    /* set table */
    std::vector<tmc::BatchSet> hs(nsets);
    std::vector<tmc::PlanSet> ps(nsets);
    uint64_t off = 0;
    double work = 0.0;
    uint32_t shells_max = 1;
    uint64_t kcap = o->max_interactions ? o->max_interactions : 0xFFFFFFFFull;
    if (kcap > 0xFFFFFFFFull)
        kcap = 0xFFFFFFFFull;
    for (uint32_t i = 0; i < nsets; ++i) {
        const tmc_params& q = sets[i].params;
        tmc::BatchSet& d = hs[i];
        std::memset(&d, 0, sizeof d);
        d.seed = sets[i].seed;
        d.first_photon = sets[i].first_photon;
        d.tally_offset = off;
        d.albedo = tmc::derived_albedo(&q);
        d.spm = tmc::derived_spm(&q);
        d.shell_max = (float)(q.shells - 1);
        d.shells = q.shells;
        d.kcap = (uint32_t)kcap;
        off += 4 * (uint64_t)q.shells;
        shells_max = std::max(shells_max, q.shells);
        ps[i] = { q.shells, q.photons, tmc::expected_interactions(&q, kcap) };
        work += ps[i].L * (double)q.photons;
        p->photons += q.photons;
    }
    p->tally_words = off;
    p->shells_max = shells_max;
    p->expected = work;

    /* launch passes and chunk queues */
    const uint32_t cus = (uint32_t)c->compute_units;
    const tmc::ShapeFn shape_of = [&](uint32_t shells, tmc::PlanShape* shape) {
        PassPlan s;
        bool fits = false;
        if (launch_shape(p, o, cus, shells, &s, &fits) != hipSuccess)
            return (int)TMC_E_HIP;
        if (!fits)
            return (int)TMC_E_INVALID; /* LDS slots of this many shells do not fit */
        *shape = { s.grid, s.block, s.wps };
        return (int)TMC_OK;
    };
    std::vector<tmc::PlanPass> passes;
    std::vector<tmc::PlanChunk> pc;
    const int prc = tmc::plan_passes(ps, o->chunk_work, TMC_BATCH_MAX_PASSES, shape_of, &passes, &pc);
    if (prc != TMC_OK) {
        plan_free(p);
        return prc;
    }
    std::vector<tmc::BatchChunk> hc(pc.size());
    for (size_t k = 0; k < pc.size(); ++k)
        hc[k] = { pc[k].set, pc[k].count, pc[k].first };
    p->npasses = (uint32_t)passes.size();
    std::vector<uint64_t> first_chunk(passes.size());
    for (size_t j = 0; j < passes.size(); ++j) {
        PassPlan& q = p->pass[j];
        q.grid = passes[j].shape.grid;
        q.block = passes[j].shape.block;
        q.wps = passes[j].shape.wps;
        q.shells_max = passes[j].hi;
        q.work = passes[j].work;
        q.chunk_work = passes[j].chunk_work;
        q.nchunks = passes[j].nchunks;
        first_chunk[j] = passes[j].first_chunk;
    }
    p->nchunks = hc.size();
    uint32_t main_pass = 0;
    for (uint32_t j = 1; j < p->npasses; ++j)
        if (p->pass[j].work > p->pass[main_pass].work)
            main_pass = j;
    if (p->npasses) {
        p->grid = p->pass[main_pass].grid;
        p->block = p->pass[main_pass].block;
        p->wps = p->pass[main_pass].wps;
        p->chunk_work = p->pass[main_pass].chunk_work;
    }

    /* device memory */
    if (hipMalloc((void**)&p->d_sets, std::max<size_t>(1, nsets) * sizeof(tmc::BatchSet)) != hipSuccess
        || hipMalloc((void**)&p->d_chunks, std::max<size_t>(1, hc.size()) * sizeof(tmc::BatchChunk)) != hipSuccess
        || hipMalloc((void**)&p->d_tally, std::max<uint64_t>(1, p->tally_words) * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&p->d_done, std::max<size_t>(1, 2 * (size_t)nsets) * sizeof(uint64_t)) != hipSuccess
        || hipMalloc((void**)&p->d_counters, TMC_BATCH_COUNTERS * sizeof(uint64_t)) != hipSuccess
        || hipHostMalloc((void**)&p->h_tally, std::max<uint64_t>(1, p->tally_words) * sizeof(uint64_t), 0)
            != hipSuccess
        || hipHostMalloc((void**)&p->h_counters, TMC_BATCH_COUNTERS * sizeof(uint64_t), 0) != hipSuccess) {
        plan_free(p);
        return TMC_E_NOMEM;
    }
    if ((nsets
         && hipMemcpyAsync(p->d_sets, hs.data(), nsets * sizeof(tmc::BatchSet), hipMemcpyHostToDevice, c->stream)
             != hipSuccess)
        || (!hc.empty()
            && hipMemcpyAsync(p->d_chunks, hc.data(), hc.size() * sizeof(tmc::BatchChunk), hipMemcpyHostToDevice,
                              c->stream)
                != hipSuccess)
        || hipStreamSynchronize(c->stream) != hipSuccess) {
        plan_free(p);
        return TMC_E_HIP;
    }

    for (uint32_t j = 0; j < p->npasses; ++j) {
        tmc::BatchArgs& a = p->pass[j].args;
        std::memset(&a, 0, sizeof a);
        a.sets = p->d_sets;
        a.chunks = p->d_chunks + first_chunk[j];
        a.nchunks = (uint32_t)p->pass[j].nchunks;
        a.shells_max = p->pass[j].shells_max;
        a.copies = copies;
        a.counters = p->d_counters;
        a.cursor = TMC_BATCH_CURSOR0 + j;
        a.tally = p->d_tally;
        a.done = p->d_done;
    }

    if (p->use_graph) {
        hipStream_t s = c->stream;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            plan_free(p);
            return TMC_E_HIP;
        }
        const hipError_t e1 = enqueue_run(p, s);
        const hipError_t e2 = hipStreamEndCapture(s, &p->graph);
        if (e1 != hipSuccess || e2 != hipSuccess
            || hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) {
            plan_free(p);
            return TMC_E_HIP;
        }
    }
    *out = p;
    return TMC_OK;
}

int plan_run_locked(tmc::Context* c, tmc_batch_plan* p, uint64_t* tallies, tmc_batch_stats* st)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (hipEventRecord(c->ev_start, c->stream) != hipSuccess)
        return TMC_E_HIP;
    if (p->use_graph) {
        if (hipGraphLaunch(p->exec, c->stream) != hipSuccess)
            return TMC_E_HIP;
    } else if (enqueue_run(p, c->stream) != hipSuccess) {
        return TMC_E_HIP;
    }
    if (hipEventRecord(c->ev_stop, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        return TMC_E_HIP;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->ev_start, c->ev_stop) != hipSuccess)
        return TMC_E_HIP;
    if (tallies)
        std::memcpy(tallies, p->h_tally, p->tally_words * sizeof(uint64_t));
    if (st) {
        st->sets = p->nsets;
        st->photons = p->photons;
        st->interactions = p->h_counters[1];
        st->truncated = p->h_counters[2];
        st->idle_lane_steps = p->h_counters[4];
        st->chunks = p->nchunks;
        st->chunk_work = p->chunk_work;
        st->expected_interactions = p->expected;
        st->kernel_ms = ms;
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st->grid = p->grid;
        st->block = p->block;
        st->copies = p->copies;
        st->waves_per_simd = p->wps;
        st->passes = p->npasses;
        st->reserved = 0;
    }
    return TMC_OK;
}

int check_sets(const tmc_set* sets, uint32_t nsets)
{
    if (!sets && nsets)
        return TMC_E_INVALID;
    for (uint32_t i = 0; i < nsets; ++i) {
        const int rc = tmc::check_params(&sets[i].params);
        if (rc != TMC_OK)
            return rc;
        if (!tmc::photon_range_ok(sets[i].first_photon, sets[i].params.photons))
            return TMC_E_RANGE;
    }
    return TMC_OK;
}

} // namespace

TMC_EXPORT void tmc_batch_options_default(tmc_batch_options* opt)
{
    if (opt)
        std::memset(opt, 0, sizeof *opt);
}

TMC_EXPORT uint64_t tmc_batch_tally_words(const tmc_set* sets, uint32_t nsets)
{
    uint64_t w = 0;
    for (uint32_t i = 0; sets && i < nsets; ++i)
        w += 4 * (uint64_t)sets[i].params.shells;
    return w;
}

TMC_EXPORT int tmc_batch_plan_create(const tmc_set* sets, uint32_t nsets, const tmc_batch_options* opt,
                                     tmc_batch_plan** plan)
{
    if (!plan)
        return TMC_E_INVALID;
    *plan = nullptr;
    int rc = check_sets(sets, nsets);
    if (rc != TMC_OK)
        return rc;
    tmc_batch_options o;
    if (opt)
        o = *opt;
    else
        tmc_batch_options_default(&o);
    std::lock_guard<std::mutex> lock(tmc::api_mutex());
    tmc::Context* c = nullptr;
    rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    return plan_create_locked(c, sets, nsets, &o, plan);
}

TMC_EXPORT int tmc_batch_plan_run(tmc_batch_plan* plan, uint64_t* tallies, tmc_batch_stats* stats)
{
    if (!plan)
        return TMC_E_INVALID;
    std::lock_guard<std::mutex> lock(tmc::api_mutex());
    tmc::Context* c = nullptr;
    const int rc = tmc::acquire_context(&c);
    if (rc != TMC_OK)
        return rc;
    return plan_run_locked(c, plan, tallies, stats);
}

TMC_EXPORT void tmc_batch_plan_destroy(tmc_batch_plan* plan)
{
    if (!plan)
        return;
    std::lock_guard<std::mutex> lock(tmc::api_mutex());
    plan_free(plan);
}

TMC_EXPORT int tmc_batch_run(const tmc_set* sets, uint32_t nsets, const tmc_batch_options* opt, uint64_t* tallies,
                             tmc_batch_stats* stats)
{
    tmc_batch_plan* plan = nullptr;
    int rc = tmc_batch_plan_create(sets, nsets, opt, &plan);
    if (rc != TMC_OK)
        return rc;
    rc = tmc_batch_plan_run(plan, tallies, stats);
    tmc_batch_plan_destroy(plan);
    return rc;
}

TMC_EXPORT int tmc_batch_autotune(const tmc_set* sets, uint32_t nsets, const tmc_batch_options* base,
                                  uint32_t sample_divisor, tmc_batch_options* best)
{
    if (!best)
        return TMC_E_INVALID;
    int rc = check_sets(sets, nsets);
    if (rc != TMC_OK)
        return rc;
    tmc_batch_options b0;
    if (base)
        b0 = *base;
    else
        tmc_batch_options_default(&b0);
    b0.use_graph = 0;
    if (sample_divisor == 0)
        sample_divisor = 1;
    auto reduced = [&](uint32_t divisor) {
        std::vector<tmc_set> r(sets, sets + nsets);
        for (auto& s : r)
            if (s.params.photons)
                s.params.photons = std::max<uint64_t>(64, s.params.photons / divisor);
        return r;
    };
    const std::vector<tmc_set> sample = reduced(sample_divisor);

    /* the automatic chunk work of the full batch on the default shape */
    tmc_batch_plan* probe = nullptr;
    rc = tmc_batch_plan_create(sets, nsets, &b0, &probe);
    if (rc != TMC_OK)
        return rc;
    const uint64_t auto_cw = probe->chunk_work;
    const double full_work = probe->expected;
    tmc_batch_plan_destroy(probe);
    /* chunk work is tuned as a multiple of the automatic value, scaled to a
     * reduced copy so that it sees the same chunk count per wave */
    auto scale_of = [&](const std::vector<tmc_set>& r, double* scale) -> int {
        tmc_batch_plan* sp = nullptr;
        const int e = tmc_batch_plan_create(r.data(), nsets, &b0, &sp);
        if (e != TMC_OK)
            return e;
        *scale = full_work > 0 ? sp->expected / full_work : 1.0;
        tmc_batch_plan_destroy(sp);
        return TMC_OK;
    };
    double scale = 1.0;
    rc = scale_of(sample, &scale);
    if (rc != TMC_OK)
        return rc;

    /* each candidate: the best of `reps` runs after a warm-up run */
    auto time_on = [&](const std::vector<tmc_set>& r, double sc, const tmc_batch_options& o, double mult,
                       int reps) -> double {
        tmc_batch_options t = o;
        t.chunk_work = (uint64_t)std::max(4096.0, mult * (double)auto_cw * sc);
        tmc_batch_plan* plan = nullptr;
        if (tmc_batch_plan_create(r.data(), nsets, &t, &plan) != TMC_OK)
            return 1e300;
        tmc_batch_stats st;
        double best_ms = 1e300;
        for (int rep = 0; rep <= reps; ++rep) /* first run warms the code object */
            if (tmc_batch_plan_run(plan, nullptr, &st) == TMC_OK && rep > 0)
                best_ms = std::min(best_ms, st.kernel_ms);
        tmc_batch_plan_destroy(plan);
        return best_ms;
    };
    auto time_of = [&](const tmc_batch_options& o, double mult) { return time_on(sample, scale, o, mult, 3); };
    /* a candidate replaces the current options only when it is faster by
     * kMargin on the sample: the sample's shorter queues favour shapes that
     * do not carry over to the full batch, and the automatic options already
     * keep nearly every lane busy */
    constexpr double kMargin = 0.97;

    tmc_batch_options cur = b0;
    double cur_mult = 1.0;
    double cur_ms = time_of(cur, cur_mult);
    for (double m : { 0.25, 4.0 }) {
        const double ms = time_of(cur, m);
        if (ms < kMargin * cur_ms) {
            cur_ms = ms;
            cur_mult = m;
        }
    }
    for (uint32_t copies : { 2u }) {
        tmc_batch_options t = cur;
        t.copies = copies;
        const double ms = time_of(t, cur_mult);
        if (ms < kMargin * cur_ms) {
            cur_ms = ms;
            cur = t;
        }
    }
    for (uint32_t block : { 64u, 512u }) {
        tmc_batch_options t = cur;
        t.block = block;
        const double ms = time_of(t, cur_mult);
        if (ms < kMargin * cur_ms) {
            cur_ms = ms;
            cur = t;
        }
    }
    for (uint32_t wps : { 6u, 4u }) {
        tmc_batch_options t = cur;
        t.waves_per_simd = wps;
        const double ms = time_of(t, cur_mult);
        if (ms < kMargin * cur_ms) {
            cur_ms = ms;
            cur = t;
        }
    }
    if (cur_ms >= 1e300)
        return TMC_E_HIP;
    /* a choice other than *base must also win on a copy four times larger
     * (the whole batch for divisors up to 4): on the 4096-set sweep the
     * 1/16 sample once picked 4x chunk work and 6 waves per SIMD, 2.4% slower
     * than the defaults on the full batch */
    const bool changed = cur_mult != 1.0 || cur.block != b0.block || cur.waves_per_simd != b0.waves_per_simd
        || cur.copies != b0.copies;
    if (changed) {
        const uint32_t div2 = sample_divisor / 4u > 1u ? sample_divisor / 4u : 1u;
        const std::vector<tmc_set> confirm = div2 > 1u ? reduced(div2) : std::vector<tmc_set>(sets, sets + nsets);
        double sc2 = 1.0;
        rc = scale_of(confirm, &sc2);
        if (rc != TMC_OK)
            return rc;
        const double base_ms = time_on(confirm, sc2, b0, 1.0, 2);
        const double pick_ms = time_on(confirm, sc2, cur, cur_mult, 2);
        if (!(pick_ms < kMargin * base_ms)) {
            cur = b0;
            cur_mult = 1.0;
        }
    }
    *best = cur;
    /* unchanged chunk work stays automatic (the plan derives it per pass) */
    best->chunk_work = cur_mult == 1.0 ? b0.chunk_work : (uint64_t)std::max(4096.0, cur_mult * (double)auto_cw);
    best->use_graph = base ? base->use_graph : 0;
    return TMC_OK;
}
