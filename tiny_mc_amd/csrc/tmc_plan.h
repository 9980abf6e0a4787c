/* Host planning shared by the entry points, with no HIP in it: parameter
 * checks and the binary32 weight cadence, the batched sweep's launch passes
 * and chunk queues, the serving tally arena and chunk sizing.  Compiled
 * into the library, and on its own into the CPU test harness that runs it
 * under the address and undefined-behaviour sanitizers. */
#ifndef TMC_PLAN_H
#define TMC_PLAN_H

#include <stdint.h>

#include <functional>
#include <map>
#include <vector>

#include "../../include/tiny_mc_gpu.h"

/* Largest weight table of the binary32 walk (entries of 16 bytes in LDS). */
#define TMC_WALK_TABLE_MAX 1024u

namespace tmc {

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

/* The binary32 weight cadence: absorptions k1 of the first cycle and k2 of
 * the second (false if the first is longer than `limit`). */
bool weight_cadence(const tmc_params* p, uint32_t limit, uint32_t* k1, uint32_t* k2);

/* Expected absorptions per photon of the FP32 weight cadence with the
 * roulette of photon.c:45-49, capped at kcap (scheduling estimate). */
double expected_interactions(const tmc_params* p, uint64_t kcap);

/* Entries of the single run's weight table for these parameters and the
 * caller's request (tmc_options.weight_table; 0: none). */
uint32_t weight_table_entries(const tmc_params* p, uint32_t precision, uint64_t photons, uint32_t request);

/* --- batched sweep ------------------------------------------------------ */

/* One set as the planner sees it: shell count, photons and the expected
 * absorptions per photon. */
struct PlanSet {
    uint32_t shells;
    uint64_t photons;
    double L;
};

/* Launch shape of a pass: grid and block, and resident waves per SIMD. */
struct PlanShape {
    uint32_t grid = 0, block = 0, wps = 0;
};

/* A run of `count` photons of set `set` from offset `first`. */
struct PlanChunk {
    uint32_t set;
    uint32_t count;
    uint64_t first;
};

/* One launch: the sets whose shell counts lie in [lo, hi], its shape, the
 * expected absorptions it carries, the absorptions per chunk, and its
 * chunks [first_chunk, first_chunk + nchunks) of the queue. */
struct PlanPass {
    uint32_t lo = 0, hi = 0;
    PlanShape shape;
    double work = 0.0;
    uint64_t chunk_work = 0;
    uint64_t first_chunk = 0, nchunks = 0;
};

/* Launch shape for LDS slots of `shells` shells: TMC_OK, TMC_E_INVALID when
 * no block fits, or another status from the device query. */
using ShapeFn = std::function<int(uint32_t shells, PlanShape* shape)>;

/* Plans a batch.  Sets are grouped by the launch shape their shell count
 * allows (a class with under a tenth of the work joins its neighbour with
 * more shells, and there are at most max_passes groups); each set is cut
 * into chunks of 64 * P photons, P = ceil(chunk_work / (64 L)), so a chunk
 * carries about chunk_work absorptions (0: a sixteenth of a wave's share of
 * its pass, at least 65536).  Within a pass the chunks are sorted longest
 * first, and the passes are ordered by their longest chunk.  `chunks`
 * receives the queues of all passes one after the other.  TMC_OK, a status
 * of shape_of, or TMC_E_RANGE when a pass holds 2^32 chunks or more. */
int plan_passes(const std::vector<PlanSet>& sets, uint64_t chunk_work, uint32_t max_passes, const ShapeFn& shape_of,
                std::vector<PlanPass>* passes, std::vector<PlanChunk>* chunks);

/* --- serving ------------------------------------------------------------ */

/* First-fit allocator of the serving tally arena (units: 64-bit words). */
struct Arena {
    std::map<uint64_t, uint64_t> free_; /* offset -> size, no two adjacent */
    void init(uint64_t words);
    bool alloc(uint64_t n, uint64_t* off);
    void release(uint64_t off, uint64_t n);
};

/* Photons of the next served chunk of a request with `left` unpublished
 * photons: about `work` absorptions (so a chunk lasts about work / 64 steps
 * of a wave, which bounds how long a new request waits for waves), between
 * 1 and `max_per_lane` photons per lane, never more than `left` (left >= 1). */
uint64_t chunk_photons(double work, double L, uint64_t max_per_lane, uint64_t left);

/* The server's chunk sizing.  Its backlog counts each request's unpublished
 * work (photons x L expected absorptions) up to `cap` (backlog_share), and
 * the chunk work and per-lane cap are the base ones times
 * serve_chunk_scale(backlog): 1 while the backlog is within
 * chunks_per_wave base chunks per resident wave, the backlog over that
 * amount beyond, at most max_scale (>= 1). */
double backlog_share(uint64_t photons, double L, double cap);
double serve_chunk_scale(double backlog, double base_work, uint32_t waves, double chunks_per_wave,
                         uint32_t max_scale);

} // namespace tmc

#endif
