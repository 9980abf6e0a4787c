// The library's host planning (tiny_mc_amd/csrc/tmc_plan.cpp) built with
// AddressSanitizer and UndefinedBehaviorSanitizer and driven over edge
// cases and seeded random inputs: parameter checks, the weight cadence and
// the table size, the batched sweep's passes and chunk queues, the serving
// arena and chunk sizing.  Besides the sanitizers' own checks it asserts the
// invariants the kernels rely on (every photon of a set in exactly one
// chunk, chunk sizes within the 32-bit fields, a set's chunks in the pass
// whose LDS slots hold its shells; arena blocks disjoint and first fit).
// Test infrastructure; prints "plan_check: ok" at the end.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "../../tiny_mc_amd/csrc/tmc_plan.h"

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "plan_check: %s:%d: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                        \
        }                                                                        \
    } while (0)

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64()
{
    // SplitMix64
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t below(uint64_t n) { return next_u64() % n; }

tmc_params params(float mu_a, float mu_s, uint32_t shells, float mps, uint64_t photons = 1)
{
    tmc_params p;
    std::memset(&p, 0, sizeof p);
    p.mu_a = mu_a;
    p.mu_s = mu_s;
    p.shells = shells;
    p.microns_per_shell = mps;
    p.photons = photons;
    return p;
}

void check_params_and_cadence()
{
    const float inf = std::numeric_limits<float>::infinity(), nan = std::numeric_limits<float>::quiet_NaN();
    CHECK(tmc::check_params(nullptr) == TMC_E_INVALID);
    const tmc_params good[] = { params(2, 20, 101, 50), params(10, 10, 1, 100), params(0.5f, 50, TMC_MAX_SHELLS, 20),
                                params(1, 0, 3, 1e6f), params(1e-4f, 100, 101, 50) };
    const tmc_params bad[] = { params(0, 20, 101, 50), params(-1, 20, 101, 50), params(2, -1, 101, 50),
                               params(2, 20, 0, 50), params(2, 20, TMC_MAX_SHELLS + 1, 50), params(2, 20, 101, 0),
                               params(nan, 20, 101, 50), params(2, inf, 101, 50), params(2, 20, 101, inf),
                               params(1e-9f, 100, 101, 50), /* albedo rounds to 1 */
                               params(2, 20, 101, 1e-45f) /* shells per mean free path overflows */ };
    for (const tmc_params& p : good)
        CHECK(tmc::check_params(&p) == TMC_OK);
    for (const tmc_params& p : bad)
        CHECK(tmc::check_params(&p) == TMC_E_INVALID);

    for (const tmc_params& p : good) {
        const float a = tmc::derived_albedo(&p);
        CHECK(a >= 0.0f && a < 1.0f && tmc::derived_spm(&p) > 0.0f);
        uint32_t k1 = 0, k2 = 0;
        if (tmc::weight_cadence(&p, 1u << 20, &k1, &k2)) {
            // the weight is still >= 0.001 after k1 - 1 absorptions and below it after k1
            float w = 1.0f;
            for (uint32_t k = 0; k + 1 < k1; ++k)
                w = w * a;
            CHECK(w >= 0.001f && w * a < 0.001f);
            CHECK(k2 >= 1 && k2 <= k1 + 1);
        } else {
            CHECK(a > 0.9999f); // only albedos this close to 1 run past 2^20 steps
        }
        const double L = tmc::expected_interactions(&p, 0);
        CHECK(std::isfinite(L) && L >= 1.0);
        CHECK(tmc::expected_interactions(&p, 5) <= 5.0);
        for (uint32_t prec = 0; prec < 4; ++prec)
            for (uint32_t req : { 0u, 1u, 700u, 5000u, TMC_WEIGHT_TABLE_OFF })
                for (uint64_t n : { (uint64_t)1, (uint64_t)1 << 20, (uint64_t)1 << 40 }) {
                    const uint32_t e = tmc::weight_table_entries(&p, prec, n, req);
                    CHECK(e <= TMC_WALK_TABLE_MAX);
                    if (prec != TMC_PRECISION_FP32 || req == TMC_WEIGHT_TABLE_OFF)
                        CHECK(e == 0);
                }
    }
    CHECK(tmc::photon_range_ok(0, 1ull << 63) && !tmc::photon_range_ok(1, 1ull << 63));
    CHECK(tmc::photon_range_ok(1ull << 63, 0) && !tmc::photon_range_ok((1ull << 63) + 1, 0));
    CHECK(!tmc::photon_range_ok(~0ull, ~0ull));
}

// A stand-in for the device occupancy query: blocks of 256 up to 2000
// shells, 128 up to 6000, 64 beyond; fewer resident workgroups as the slots
// grow; nothing fits beyond `limit` shells.
int fake_shape(uint32_t shells, tmc::PlanShape* s, uint32_t limit)
{
    if (shells > limit)
        return TMC_E_INVALID;
    const uint32_t block = shells <= 2000 ? 256 : shells <= 6000 ? 128 : 64;
    const uint32_t n = shells <= 500 ? 4 : shells <= 2000 ? 2 : 1;
    s->grid = 256 * n;
    s->block = block;
    s->wps = n * (block / 64) / 4;
    return TMC_OK;
}

void check_plan(const std::vector<tmc::PlanSet>& sets, uint64_t chunk_work, uint32_t max_passes)
{
    std::vector<tmc::PlanPass> passes;
    std::vector<tmc::PlanChunk> chunks;
    const int rc = tmc::plan_passes(
        sets, chunk_work, max_passes, [](uint32_t sh, tmc::PlanShape* s) { return fake_shape(sh, s, 1u << 20); },
        &passes, &chunks);
    CHECK(rc == TMC_OK);
    CHECK(passes.size() <= max_passes);
    CHECK(sets.empty() == passes.empty());
    // passes partition the queue
    uint64_t at = 0;
    for (const tmc::PlanPass& q : passes) {
        CHECK(q.first_chunk == at);
        at += q.nchunks;
        CHECK(q.lo <= q.hi && q.chunk_work >= 1 && q.shape.grid > 0 && q.shape.block % 64 == 0);
        // longest chunk first within a pass
        double prev = 1e300;
        for (uint64_t k = q.first_chunk; k < q.first_chunk + q.nchunks; ++k) {
            const tmc::PlanChunk& c = chunks[k];
            CHECK(c.set < sets.size());
            CHECK(sets[c.set].shells >= q.lo && sets[c.set].shells <= q.hi);
            const double dur = (double)((c.count + 63) / 64) * sets[c.set].L;
            CHECK(dur <= prev);
            prev = dur;
        }
    }
    CHECK(at == chunks.size());
    // every set's shells fall in exactly one pass, and its photons are cut
    // into contiguous chunks that cover [0, N) once
    std::vector<std::vector<const tmc::PlanChunk*>> per(sets.size());
    for (const tmc::PlanChunk& c : chunks)
        per[c.set].push_back(&c);
    for (size_t i = 0; i < sets.size(); ++i) {
        size_t owners = 0;
        for (const tmc::PlanPass& q : passes)
            owners += sets[i].shells >= q.lo && sets[i].shells <= q.hi;
        CHECK(owners == 1);
        uint64_t next = 0, C = 0;
        for (size_t k = 0; k < per[i].size(); ++k) {
            const tmc::PlanChunk* c = per[i][k];
            CHECK(c->first == next && c->count >= 1 && c->count <= (1u << 31));
            if (k == 0)
                C = c->count;
            else
                CHECK(c->count <= C); // all chunks but the last are full
            if (k + 1 < per[i].size())
                CHECK(c->count == C && C % 64 == 0);
            next += c->count;
        }
        CHECK(next == sets[i].photons);
        CHECK(per[i].size() <= (1u << 24) + 1);
    }
}

void check_plans()
{
    // empty, one set, zero photons
    check_plan({}, 0, 3);
    check_plan({ { 101, 0, 72.0 } }, 0, 3);
    check_plan({ { 101, 1, 72.0 } }, 0, 3);
    check_plan({ { 1, 63, 1.0 }, { 1, 64, 1.0 }, { 1, 65, 1.0 } }, 1, 3);
    // many photons per set in few chunks
    check_plan({ { 101, 1ull << 34, 72.0 } }, 1ull << 30, 3);
    // the cap of 2^24 photons per lane
    check_plan({ { 101, 1ull << 33, 1e-9 } }, 1ull << 40, 3);
    // seeded random sweeps over shells, photons and work per photon
    for (int trial = 0; trial < 60; ++trial) {
        std::vector<tmc::PlanSet> sets(1 + below(40));
        for (tmc::PlanSet& s : sets) {
            const uint32_t cls = (uint32_t)below(4);
            s.shells = cls == 0 ? 1 + (uint32_t)below(500) : cls == 1 ? 501 + (uint32_t)below(1500)
                : cls == 2 ? 2001 + (uint32_t)below(4000) : 6001 + (uint32_t)below(TMC_MAX_SHELLS - 6000);
            s.photons = below(3) == 0 ? below(100) : below(1ull << (8 + below(13)));
            s.L = 1.0 + (double)below(20000) / 10.0;
        }
        check_plan(sets, below(2) ? 0 : 1 + below(1u << 20), 1 + (uint32_t)below(3));
    }
    // a shell count that no block fits is refused with the shape's status
    std::vector<tmc::PlanPass> passes;
    std::vector<tmc::PlanChunk> chunks;
    const int rc = tmc::plan_passes(
        { { 100, 10, 2.0 }, { 900, 10, 2.0 } }, 0, 3,
        [](uint32_t sh, tmc::PlanShape* s) { return fake_shape(sh, s, 500); }, &passes, &chunks);
    CHECK(rc == TMC_E_INVALID);
}

void check_arena()
{
    const uint64_t W = 1024;
    tmc::Arena a;
    a.init(W);
    std::vector<int> owner(W, -1);
    struct Block {
        uint64_t off, n;
    };
    std::vector<Block> live;
    for (int step = 0; step < 6000; ++step) {
        if (live.empty() || below(3) != 0) {
            const uint64_t n = 1 + below(below(8) == 0 ? 256 : 32);
            // first fit: the lowest offset with n free words in a row
            uint64_t expect = W, run = 0;
            for (uint64_t i = 0; i < W; ++i) {
                run = owner[i] < 0 ? run + 1 : 0;
                if (run == n) {
                    expect = i + 1 - n;
                    break;
                }
            }
            uint64_t off = 0;
            const bool ok = a.alloc(n, &off);
            CHECK(ok == (expect < W));
            if (!ok)
                continue;
            CHECK(off == expect && off + n <= W);
            for (uint64_t i = off; i < off + n; ++i) {
                CHECK(owner[i] < 0);
                owner[i] = step;
            }
            live.push_back({ off, n });
        } else {
            const size_t k = below(live.size());
            const Block b = live[k];
            live[k] = live.back();
            live.pop_back();
            for (uint64_t i = b.off; i < b.off + b.n; ++i)
                owner[i] = -1;
            a.release(b.off, b.n);
        }
        // the free list is exactly the free words, in maximal runs
        uint64_t free_words = 0, last_end = ~0ull;
        for (const auto& kv : a.free_) {
            CHECK(kv.second > 0 && kv.first + kv.second <= W && kv.first != last_end);
            for (uint64_t i = kv.first; i < kv.first + kv.second; ++i)
                CHECK(owner[i] < 0);
            free_words += kv.second;
            last_end = kv.first + kv.second;
        }
        uint64_t unowned = 0;
        for (int o : owner)
            unowned += o < 0;
        CHECK(free_words == unowned);
    }
    tmc::Arena e;
    e.init(0);
    uint64_t off = 0;
    CHECK(!e.alloc(1, &off));
}

void check_chunk_photons()
{
    for (int t = 0; t < 100000; ++t) {
        const double work = (double)below(1ull << 30);
        const double L = 1.0 + (double)below(100000) / 7.0;
        const uint64_t mpl = 1 + below(1u << 20);
        const uint64_t left = 1 + below(below(2) ? 100 : 1ull << 40);
        const uint64_t r = tmc::chunk_photons(work, L, mpl, left);
        CHECK(r >= 1 && r <= left && r <= 64 * mpl);
        CHECK(r == left || r % 64 == 0);
    }
}

/* The server's chunk sizing: the scale is 1 up to the threshold
 * (chunks_per_wave base chunks per wave), the backlog over the threshold
 * beyond, at most max_scale, and monotone; one request's share is capped,
 * so a request alone raises the scale to at most cap / threshold while many
 * reach max_scale; and publishing a request chunk by chunk as the server
 * does, each step taking share(before) - share(after) off the backlog,
 * brings the backlog back to zero within rounding once the request is
 * published (share(0) = 0, no drift over many chunks). */
void check_chunk_scale()
{
    for (int t = 0; t < 20000; ++t) {
        const double base = (double)(1 + below(1u << 20));
        const uint32_t waves = 1 + (uint32_t)below(1u << 14);
        const double cpw = 1.0 + (double)below(4);
        const uint32_t max_scale = 1 + (uint32_t)below(1024);
        const double thr = cpw * base * (double)waves;
        const double b = thr * (double)below(1u << 12) / 256.0;
        const double sc = tmc::serve_chunk_scale(b, base, waves, cpw, max_scale);
        CHECK(sc >= 1.0 && sc <= (double)max_scale);
        if (b <= thr)
            CHECK(sc == 1.0);
        else if (b / thr <= (double)max_scale)
            CHECK(sc == b / thr);
        else
            CHECK(sc == (double)max_scale);
        CHECK(tmc::serve_chunk_scale(b * 1.5 + 1.0, base, waves, cpw, max_scale) >= sc);
    }
    const double base = 131072.0, cpw = 2.0;
    const uint32_t waves = 6144, max_scale = 128;
    const double thr = cpw * base * waves, cap = 2.0 * thr;
    /* one request of 2^24 photons at 5,750 absorptions each (1e11) alone */
    const double one = tmc::backlog_share(1ull << 24, 5750.0, cap);
    CHECK(one == cap);
    CHECK(tmc::serve_chunk_scale(one, base, waves, cpw, max_scale) == 2.0);
    CHECK(tmc::backlog_share(1000, 75.5, cap) == 1000 * 75.5);
    /* 200 such requests queued reach the largest scale */
    CHECK(tmc::serve_chunk_scale(200 * one, base, waves, cpw, max_scale) == (double)max_scale);
    for (int t = 0; t < 2000; ++t) {
        const uint64_t N = 1 + below(below(2) ? 1ull << 24 : 4096);
        const double L = 1.0 + (double)below(20000) / 3.0;
        const double c = (double)(1 + below(1ull << 34));
        double backlog = tmc::backlog_share(N, L, c);
        const double added = backlog;
        uint64_t next = 0;
        while (next < N) {
            const uint64_t cnt = std::min<uint64_t>(N - next, 1 + below(1 + N / 8));
            backlog -= tmc::backlog_share(N - next, L, c) - tmc::backlog_share(N - next - cnt, L, c);
            next += cnt;
        }
        CHECK(std::fabs(backlog) <= 1e-9 * added + 1e-6);
    }
}

} // namespace

int main()
{
    check_params_and_cadence();
    check_plans();
    check_arena();
    check_chunk_photons();
    check_chunk_scale();
    std::puts("plan_check: ok");
    return 0;
}
