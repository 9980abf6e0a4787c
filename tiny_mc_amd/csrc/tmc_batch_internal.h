/* Device-side layout of a batched run (set table, chunk queue) shared by the
 * multi-set kernel (tmc_batch.hip) and its host side (tmc_batch_api.cpp). */
#ifndef TMC_BATCH_INTERNAL_H
#define TMC_BATCH_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmc {

/* One parameter set, as the kernel reads it (64 bytes). */
struct BatchSet {
    uint64_t seed; /* Philox key */
    uint64_t first_photon; /* stream index of the set's photon 0 */
    uint64_t tally_offset; /* word offset of the set's 4*shells tally words */
    float albedo; /* binary32, photon.c:8 */
    float spm; /* binary32, photon.c:9 */
    float shell_max; /* (float)(shells - 1) */
    uint32_t shells;
    uint32_t kcap; /* stop a photon after kcap absorptions */
    uint32_t pad;
    uint64_t photons; /* serving: photons of the request */
    uint64_t token; /* serving: written to the slot's flag when the request is done */
};
static_assert(sizeof(BatchSet) == 64, "BatchSet layout");

/* A contiguous photon range of one set (16 bytes). */
struct BatchChunk {
    uint32_t set;
    uint32_t count; /* photons, >= 1 */
    uint64_t first; /* offset within the set */
};
static_assert(sizeof(BatchChunk) == 16, "BatchChunk layout");

/* Serving: one entry of the chunk ring (32 bytes, host-coherent memory).
 * The host writes set, count and first, then stores seq = the entry's ring
 * index t with release; a wave holding ticket t takes the entry only when
 * it reads seq == t (acquire), and after reading it stores seq = t |
 * TMC_RING_CONSUMED (release).  The host rewrites the slot for index t +
 * ring only once it reads that mark, so no entry is overwritten before its
 * wave has read it, whatever the timing.
 * count == 0 marks a cancellation: `first` photons of request `set` that
 * will never be published; the wave adds them to the request's completed
 * photons, so the request finishes (and frees its slot) once the chunks
 * already published are done. */
struct RingEntry {
    uint32_t set; /* slot | generation << TMC_SERVE_SLOT_BITS */
    uint32_t count; /* photons, or 0: cancellation of `first` photons */
    uint64_t first; /* offset within the request, or the cancelled photons */
    uint64_t seq; /* ring index of the entry, | TMC_RING_CONSUMED once read */
    uint64_t pad;
};
static_assert(sizeof(RingEntry) == 32, "RingEntry layout");

#define TMC_RING_CONSUMED (1ull << 63)

/* Serving: control words the host writes in host-coherent memory. */
struct ServeControl {
    uint64_t tail; /* chunks published so far (ring indices below it are valid) */
    uint64_t stop; /* launch id of the kernel asked to leave */
    uint64_t claimed; /* written by the kernel: one past the highest ring index claimed */
    uint64_t pad[5];
};

/* Serving: device-memory copies of the control words, each on its own
 * 64-byte line, refreshed from the host by one wave at a time, at most once
 * per refresh interval, so the polling of thousands of waves stays off the
 * host bus.  Wave w polls copy w % TMC_SERVE_CACHE_COPIES: with a single
 * line, thousands of waves without work polled it every few microseconds
 * and the waves still walking waited behind that traffic on every poll.
 * The line after the copies holds the time of the last refresh (its tail
 * word), read only by the waves of copy 0, which elect the refresher. */
struct ServeCache {
    uint64_t tail;
    uint64_t stop;
    uint64_t pad[6];
};
static_assert(sizeof(ServeCache) == 64, "ServeCache layout");
#define TMC_SERVE_CACHE_COPIES 64u

struct BatchArgs {
    const BatchSet* sets; /* batch: device table; serve: host-coherent table */
    const BatchChunk* chunks; /* batch: device queue */
    RingEntry* ring; /* serve: host-coherent chunk ring */
    uint32_t nchunks; /* batch: queue length of this pass */
    uint32_t cursor; /* batch: index of this pass's chunk cursor in counters */
    uint32_t shells_max; /* LDS slot size in shells */
    uint32_t copies; /* tally copies per wave (1 or 2) */
    uint32_t ring_mask; /* serve: ring size - 1 (power of two) */
    unsigned long long* counters; /* [0] serve: tickets of this launch [1] interactions [2] truncated
                                     [3] serve: waves that have left [4] batch: idle lane-steps
                                     [5 + j] batch: chunk cursor of pass j (TMC_BATCH_COUNTERS words) */
    unsigned long long* tally; /* tally pool */
    unsigned long long* done; /* per set: [2s] photons completed and flushed, [2s+1] absorptions */
    /* serving only */
    ServeControl* hctl; /* host-coherent control words */
    ServeCache* dctl; /* device copies of the control words (TMC_SERVE_CACHE_COPIES + 1 lines) */
    unsigned long long* out; /* host-coherent mirror of the tally pool: sums of finished requests */
    unsigned long long* flags; /* host-coherent, per slot: [2s] token of the finished request, [2s+1]
                                  its absorptions */
    uint64_t ticket_base; /* ring index of this launch's first ticket */
    uint64_t launch_id; /* the stop word that makes this launch leave */
    uint32_t total_waves; /* waves of the launch */
    uint32_t refresh_ticks; /* realtime ticks (100 MHz) between reads of the host control words */
};

/* Deposits between flushes of a wave's LDS slot: the wave flushes once its
 * absorptions since the last flush reach this count, so a slot word takes
 * fewer than 2^20 + 64 deposits (a wave makes at most 64 per step) of at
 * most 2^43 each (the heat2 image of w = 1): below 2^64.  A constant rather
 * than a kernel argument: the walk loop keeps no register for it. */
#define TMC_SLOT_FLUSH_DEPOSITS (1ull << 20)
/* Most wave steps between two flushes (a step deposits at least once): the
 * horizon of the absorption-cap test. */
#define TMC_SLOT_FLUSH_STEPS (uint32_t)TMC_SLOT_FLUSH_DEPOSITS

/* 64-bit counter words of a batched launch. */
#define TMC_BATCH_COUNTERS 8
/* Launches of a batch, each over the sets of one LDS-slot size class. */
#define TMC_BATCH_MAX_PASSES 3
#define TMC_BATCH_CURSOR0 5

/* Serving: a set field of a chunk is (slot | generation << 16), so a slot
 * reused by a later request never matches a wave's stale set. */
#define TMC_SERVE_SLOT_BITS 16u

size_t batch_lds_bytes(uint32_t block, uint32_t shells_max, uint32_t copies);

hipError_t launch_batch(uint32_t sampler, uint32_t math, const BatchArgs& a, uint32_t grid, uint32_t block,
                        hipStream_t stream);
hipError_t launch_batch_clear(unsigned long long* tally, uint64_t tally_words, unsigned long long* done,
                              uint32_t nsets, unsigned long long* counters, hipStream_t stream);
hipError_t launch_serve(uint32_t sampler, uint32_t math, const BatchArgs& a, uint32_t grid, uint32_t block,
                        hipStream_t stream);
hipError_t serve_occupancy(uint32_t sampler, uint32_t math, uint32_t block, uint32_t shells_max, uint32_t copies,
                           int* wg_per_cu);
hipError_t batch_occupancy(uint32_t sampler, uint32_t math, uint32_t block, uint32_t shells_max, uint32_t copies,
                           int* wg_per_cu);

} // namespace tmc

#endif
