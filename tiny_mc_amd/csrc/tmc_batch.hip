/* Many parameter sets in one persistent launch (batched sweep).
 *
 * The host splits every set (mu_a, mu_s, shells, width, photon budget, seed)
 * into chunks of photons with about equal expected work, sorts the chunks
 * longest first, and uploads the set table and the chunk queue once.  Each
 * wave claims the next chunk from a global cursor and walks its photons with
 * per-lane regeneration; the set is uniform across the wave, so the set's
 * physics and Philox key stay in scalar registers.  A wave moves to a chunk
 * of another set only when all its lanes are dead (the drain), which the
 * chunk sizes keep short relative to the chunk.
 *
 * Tallies: every wave owns an LDS slot of shells_max shells (1 or 2 copies,
 * lanes 0-31 / 32-63 in separate copies to halve same-address collisions),
 * flushed into the set's 128-bit global sums when the wave changes set,
 * every TMC_SLOT_FLUSH_DEPOSITS absorptions, and at exit.  With each flush
 * the wave adds the photons it completed for that set to the set's done
 * counter (serving: once its sum updates have completed), so done ==
 * photons means the set's tallies are complete.
 *
 * The per-absorption arithmetic is the same step<>() as the single-run
 * kernel, so each set's tallies are bit-identical to a single run of that
 * set (the stream is keyed by (seed, photon index) only).
 *
 * Serving (SERVE = true) runs the same loop as a resident kernel fed by a
 * server thread through host-coherent memory only, so the server never
 * waits on a copy queued behind some other kernel:
 *   - the request table, the chunk ring and the control words (tail, stop)
 *     live in host memory; waves read them with system-scope loads, the
 *     tail through a device-memory cache that one wave at a time refreshes;
 *   - a wave whose lanes have drained takes a ticket (ring index) and keeps
 *     it until the tail passes it; a wave leaves only when the host has
 *     asked this launch to stop,
 *     which the host does only when every published chunk is done;
 *   - the wave whose completion count reaches a request's photons moves
 *     the request's sums (and absorptions) into a host mirror, zeroes them
 *     on the device for the slot's next request, and then writes the
 *     request's token into the slot's host flag.
 */
#include "tmc_batch_internal.h"
#include "tmc_device.h"
#include "tmc_step.h"

namespace tmc {

/* Copies of the overflow shell's sums in each copy of a wave's slot
 * (lane & 15 picks one), after the inner shells (LaneTally).  Shells and
 * copies alike are (w image, w*w image) pairs of adjacent words. */
#ifndef TMC_SLOT_HOT
#define TMC_SLOT_HOT 16u
#endif
static_assert(TMC_SLOT_HOT <= TMC_CHECK_HOT, "the shell check covers every overflow copy");

/* 64-bit words of one copy of a slot: room for the inner shells of a set of
 * up to shells_max shells and its overflow copies. */
__host__ __device__ __forceinline__ size_t slot_copy_words(uint32_t shells_max)
{
    return 2u * ((size_t)shells_max - 1u + TMC_SLOT_HOT);
}

/* Dead lanes a wave gathers before it launches photons into them (or
 * until it has drained).  Measured on the 4096-set sweep: 1 -> 12.24 s,
 * 2 -> 12.20, 4 -> 12.21.  (The single-run walk's scalar lane mask with
 * every lane stepping measured 3% slower here: 12.66 s.) */
#ifndef TMC_SLOT_RELAUNCH_MIN
#define TMC_SLOT_RELAUNCH_MIN 2
#endif

/* 64-bit words of one wave's LDS slot. */
__host__ __device__ __forceinline__ size_t slot_words(uint32_t shells_max, uint32_t copies)
{
    return (size_t)copies * slot_copy_words(shells_max);
}

/* Moves one wave's private LDS slot into a set's global sums and zeroes it.
 * Only this wave touches the slot, and a wave's LDS operations complete in
 * order, so plain loads see every earlier ds_add of the wave.  Serving uses
 * system-scope atomics, so the wave that finishes the request reads every
 * wave's additions when it takes the sums. */
template <bool SERVE>
__device__ __forceinline__ void flush_slot(unsigned long long* slot, uint32_t shells_max, uint32_t copies,
                                           uint32_t shells, unsigned long long* g, uint32_t lane)
{
    const uint32_t last = shells - 1;
    for (uint32_t q = lane; q < last + TMC_SLOT_HOT; q += 64) {
        unsigned long long a = 0, b = 0;
        for (uint32_t c = 0; c < copies; ++c) {
            unsigned long long* lw = slot + (size_t)c * slot_copy_words(shells_max);
            a += lw[2 * q];
            b += lw[2 * q + 1];
            lw[2 * q] = 0;
            lw[2 * q + 1] = 0;
        }
        const uint32_t shell = q < last ? q : last;
        if (SERVE) {
            if (a)
                add128_sys(g + 2 * shell, a);
            if (b)
                add128_sys(g + 2 * shells + 2 * shell, b);
        } else {
            if (a)
                add128(g + 2 * shell, a);
            if (b)
                add128(g + 2 * shells + 2 * shell, b);
        }
    }
}

/* Uniform view of the set a wave is walking. */
struct SetState {
    Physics phys;
    uint32_t shells;
    uint32_t kcap;
    uint64_t first_photon;
    unsigned long long* tally;
    unsigned long long* done;
    /* serving */
    uint64_t tally_offset;
    uint64_t photons;
    uint64_t token;
    uint32_t slot;
};

__device__ __forceinline__ uint64_t load_sys(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void store_sys(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

/* Reads a word at the point where the system-scope atomics are performed
 * and leaves zero behind. */
__device__ __forceinline__ unsigned long long take_sys(unsigned long long* p)
{
    return __hip_atomic_exchange(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

/* The kernel's arguments where a rare path (a claim, a flush, a set switch)
 * uses them.  Serving: read through the argument block's pointer passed
 * through an empty asm, so each rare path loads the fields it needs (scalar
 * loads, cached) and no load is hoisted out of it: the serving kernel's loop
 * then holds no argument in scalar registers, which leaves room for the
 * set's Philox round keys.  Batch: the arguments as passed.  The argument
 * block starts with the BatchArgs the kernel takes by value (explicit
 * kernel arguments are laid out from offset 0 in declaration order), so
 * walk_multi must keep it as its first parameter. */
template <bool SERVE>
__device__ __forceinline__ const BatchArgs& args_at_use(const BatchArgs& a)
{
    if (!SERVE)
        return a;
    typedef const __attribute__((address_space(4))) BatchArgs* ConstArgs;
    ConstArgs p = (ConstArgs)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const BatchArgs*)p;
}

/* Loads set `index` into the wave-uniform state, with the set's Philox
 * round keys formed once and pinned to scalar registers (formed inside the
 * step instead, they cost 18 scalar additions per absorption).  Batch
 * tables are device memory written before the launch; serving tables are
 * written by the host while the kernel runs, so lane 0 reads them with
 * system-scope loads and broadcasts the words. */
template <uint32_t SAMPLER, bool SERVE>
__device__ __forceinline__ void load_set(SetState& st, const BatchArgs& args, uint32_t index, uint32_t lane)
{
    const BatchArgs& a = args_at_use<SERVE>(args);
    /* words: seed, first_photon, tally_offset, (albedo | spm << 32),
     * (shell_max | shells << 32), (kcap | pad << 32), photons, token */
    constexpr int NW = SERVE ? 8 : 6;
    uint64_t w[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
    const uint64_t* src = (const uint64_t*)(a.sets + index);
    if (SERVE) {
        uint64_t v[NW];
        for (int i = 0; i < NW; ++i)
            v[i] = 0;
        if (lane == 0)
            for (int i = 0; i < NW; ++i)
                v[i] = load_sys(src + i);
        for (int i = 0; i < NW; ++i)
            w[i] = rfl64(v[i]);
    } else {
        for (int i = 0; i < NW; ++i)
            w[i] = src[i];
    }
    st.phys.key0 = (uint32_t)w[0];
    st.phys.key1 = (uint32_t)(w[0] >> 32);
    st.first_photon = rfl64(w[1]); /* scalar: the base of every photon launch */
    st.tally_offset = w[2];
    st.tally = a.tally + w[2];
    st.phys.albedo = __uint_as_float((uint32_t)w[3]);
    st.phys.spm = __uint_as_float((uint32_t)(w[3] >> 32));
    st.phys.shell_max = __uint_as_float((uint32_t)w[4]);
    st.shells = (uint32_t)(w[4] >> 32);
    set_derived<true>(st.phys);
    st.kcap = (uint32_t)w[5];
    st.done = a.done + 2 * (uint64_t)index;
    st.photons = w[6];
    st.token = w[7];
    st.slot = index;
}

enum ClaimResult { CLAIM_GOT, CLAIM_EMPTY, CLAIM_DONE };

/* Tries to give the wave its next chunk (wave-uniform result).
 * Batch: the next index of the fixed queue, or DONE past its end.
 * Serving: the wave (drained) takes a ticket (the next ring index) once and
 * keeps it until its entry is published; meanwhile the call returns EMPTY
 * without waiting, and the wave backs off.  The tail is read from
 * the wave's copy of the control words (`group`); when the ticket lies past
 * it and the last refresh is older than the refresh interval, the one wave
 * of group 0 that wins the refresh reads the host's words (stop first, with
 * acquire: a stop seen means the tail read after it is final) and its lanes
 * write them into every copy.  Below the tail the entry is taken once its
 * sequence number equals the ticket (RingEntry).  DONE once this launch is
 * asked to stop and the ticket lies past the tail: the host asks only when
 * every published entry was taken, so an entry at the ticket can only come
 * after the launch has ended. */
template <bool SERVE>
__device__ __forceinline__ ClaimResult try_claim(const BatchArgs& args, uint32_t lane, uint32_t group,
                                                 uint64_t& ticket, BatchChunk& ch)
{
    const BatchArgs& a = args_at_use<SERVE>(args);
    if (!SERVE) {
        unsigned long long c = 0;
        if (lane == 0)
            c = atomicAdd(a.counters + a.cursor, 1ull);
        c = rfl64(c);
        if (c >= a.nchunks)
            return CLAIM_DONE;
        ch = a.chunks[c];
        return CLAIM_GOT;
    }
    ServeCache* const copy = a.dctl + group;
    uint64_t t = ticket, tail = 0, w0 = 0, w1 = 0;
    uint32_t res = CLAIM_EMPTY, refresh = 0;
    if (lane == 0) {
        if (t == ~0ull)
            t = a.ticket_base + atomicAdd(a.counters, 1ull);
        /* the cached tail only tells when to look: the entry's own sequence
         * number decides, so relaxed loads suffice here */
        tail = load_sys(&copy->tail);
        if (t >= tail && group == 0) {
            uint64_t* const stamp = &a.dctl[TMC_SERVE_CACHE_COPIES].tail;
            const uint64_t now = __builtin_amdgcn_s_memrealtime();
            uint64_t last = load_sys(stamp);
            refresh = now - last >= a.refresh_ticks
                && __hip_atomic_compare_exchange_strong(stamp, &last, now, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    uint64_t stop = 0;
    const bool refreshed = rfl32(refresh) != 0;
    if (refreshed) {
        uint64_t ht = 0;
        if (lane == 0) {
            stop = __hip_atomic_load(&a.hctl->stop, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            ht = load_sys(&a.hctl->tail);
        }
        stop = rfl64(stop);
        ht = rfl64(ht);
        if (lane < TMC_SERVE_CACHE_COPIES) {
            __hip_atomic_fetch_max(&a.dctl[lane].tail, ht, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_exchange(&a.dctl[lane].stop, stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        tail = ht > tail ? ht : tail;
    }
    if (lane == 0) {
        if (t < tail) {
            RingEntry* e = a.ring + (t & a.ring_mask);
            /* the host stores seq last, with release: seq == t (acquire)
             * means the entry's fields are the ones published for t */
            if (__hip_atomic_load(&e->seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) == t) {
                w0 = load_sys((const uint64_t*)e);
                w1 = load_sys(&e->first);
                /* consumed (release: after the field reads): the host may
                 * now reuse the slot */
                __hip_atomic_store(&e->seq, t | TMC_RING_CONSUMED, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                res = CLAIM_GOT;
                /* progress report for the host's backlog estimate (a stale
                 * value only makes the host publish less) */
                store_sys(&a.hctl->claimed, t + 1);
            }
        } else {
            if (!refreshed)
                stop = load_sys(&copy->stop);
            if (stop == a.launch_id)
                res = CLAIM_DONE;
        }
    }
    ticket = rfl64(t);
    res = rfl32(res);
    if (res == CLAIM_GOT) {
        ticket = ~0ull;
        w0 = rfl64(w0);
        w1 = rfl64(w1);
        ch.set = (uint32_t)w0;
        ch.count = (uint32_t)(w0 >> 32);
        ch.first = w1;
    }
    return (ClaimResult)res;
}

/* Serving: called by the wave whose publication completed the request.
 * Every other contribution to its sums and counters was performed before
 * the completing count (each wave waits for its sum updates before it
 * counts), so exchanging each word with zero reads the final sums and
 * leaves the region clear for the slot's next request.  The sums go to
 * the host mirror, then, once those writes have completed, the token goes
 * to the slot's flag. */
__device__ __forceinline__ void finish_request(const BatchArgs& args, const SetState& st, uint32_t lane)
{
    const BatchArgs& a = args_at_use<true>(args);
    const uint32_t words = 4 * st.shells;
    for (uint32_t i = lane; i < words; i += 64)
        store_sys((uint64_t*)a.out + st.tally_offset + i, take_sys(st.tally + i));
    if (lane == 0) {
        const unsigned long long absorbed = take_sys(st.done + 1);
        take_sys(st.done);
        store_sys((uint64_t*)a.flags + 2 * (uint64_t)st.slot + 1, absorbed);
    }
    /* every lane's mirror stores before lane 0's flag (the release below
     * orders only lane 0's own) */
    wait_vmem();
    __builtin_amdgcn_wave_barrier();
    if (lane == 0)
        __hip_atomic_store((uint64_t*)a.flags + 2 * (uint64_t)st.slot, st.token, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

/* Serving: counts `photons` of the request as completed without walking
 * them (a cancellation's photons that were never published); the wave whose
 * count completes the request finishes it. */
__device__ __forceinline__ void publish_skip(const BatchArgs& a, const SetState& st, uint32_t lane, uint64_t photons)
{
    uint32_t last = 0;
    if (lane == 0) {
        const unsigned long long old = __hip_atomic_fetch_add(st.done, (unsigned long long)photons, __ATOMIC_ACQ_REL,
                                                              __HIP_MEMORY_SCOPE_AGENT);
        last = old + photons == st.photons;
    }
    if (rfl32(last))
        finish_request(a, st, lane);
}

/* Flushes the slot into the set's sums, then publishes the photons completed
 * since the last publication (and their absorptions) to the set, and the
 * absorptions and idle lane-steps (batch) of the steps since then to the
 * launch counters (published here rather than summed per step, to keep the
 * walk loop's scalar registers few).  A wave publishes at every flush, so
 * between two publications it makes fewer than TMC_SLOT_FLUSH_DEPOSITS + 64
 * absorptions, as many steps at most, and completes at most as many photons
 * as it launched plus 64: 32-bit counts.  Serving: the sums and
 * the counters are system-scope atomics, and the wave waits for every sum
 * update to complete before it adds to the counters, so the wave whose
 * count completes the request finds final sums. */
template <bool SERVE>
__device__ __forceinline__ void flush_and_publish(unsigned long long* slot, const BatchArgs& args, const SetState& st,
                                                  uint32_t lane, uint32_t& completed, uint32_t& absorbed,
                                                  uint32_t& steps)
{
    const BatchArgs& a = args_at_use<SERVE>(args);
    flush_slot<SERVE>(slot, a.shells_max, a.copies, st.shells, st.tally, lane);
    if (!completed && !absorbed)
        return;
    if (lane == 0) {
        atomicAdd(a.counters + 1, (unsigned long long)absorbed);
        if (!SERVE)
            atomicAdd(a.counters + 4, (unsigned long long)steps * 64u - absorbed);
    }
    steps = 0;
    if (SERVE) {
        /* every lane's sum updates before lane 0's count (the release below
         * orders only lane 0's own) */
        wait_vmem();
        __builtin_amdgcn_wave_barrier();
        uint32_t last = 0;
        if (lane == 0) {
            __hip_atomic_fetch_add(st.done + 1, (unsigned long long)absorbed, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
            /* release: the sums and absorptions before the count; acquire:
             * the wave whose count completes the request sees every other
             * wave's sums (each counted after its own) */
            const unsigned long long old = __hip_atomic_fetch_add(st.done, (unsigned long long)completed,
                                                                  __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            last = completed && old + completed == st.photons;
        }
        if (rfl32(last))
            finish_request(a, st, lane);
    } else if (lane == 0) {
        atomicAdd(st.done, (unsigned long long)completed);
        atomicAdd(st.done + 1, (unsigned long long)absorbed);
    }
    completed = 0;
    absorbed = 0;
}

template <uint32_t SAMPLER, bool SERVE, uint32_t MATH>
__global__ void __launch_bounds__(1024) walk_multi(BatchArgs a)
{
    extern __shared__ unsigned long long lds_slots[];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    unsigned long long* slot = lds_slots + (size_t)wave * slot_words(a.shells_max, a.copies);
    for (uint32_t i = lane; i < slot_words(a.shells_max, a.copies); i += 64)
        slot[i] = 0;
    /* fmax is set with each set (its shell_max) */
    LaneTally lt = { slot + (a.copies == 2 && lane >= 32 ? slot_copy_words(a.shells_max) : 0), 0.0f, ~0ull };
    const uint32_t slot_mask = SERVE ? (1u << TMC_SERVE_SLOT_BITS) - 1u : 0xFFFFFFFFu;

    SetState st;
    uint32_t cur_set = 0xFFFFFFFFu;
    uint64_t pool_next = 0, pool_end = 0; /* photon offsets within cur_set */
    /* wave flags, one scalar word */
    enum : uint32_t { PENDING = 1u, QUEUE_DONE = 2u, CAPPING = 4u /* per-step cap checks until the next flush */ };
    uint32_t flags = 0;
    BatchChunk pend = { 0u, 0u, 0ull };
    /* absorptions, wave steps and photons launched since the last
     * publication (32-bit: see flush_and_publish), and the lanes live at it:
     * the photons completed in between are launched + live_pub - live now,
     * formed only when they are published (a per-step count of the lanes
     * whose photon ended cost scalar instructions on every step) */
    uint32_t absorbed = 0, steps = 0, launched = 0, live_pub = 0;
    uint64_t ticket = ~0ull; /* serving: ring index held, ~0 if none */
    /* serving: failed polls in a row of a dry wave (its back-off) */
    uint32_t fails = 0;
    const uint32_t group = SERVE ? (blockIdx.x * (blockDim.x >> 6) + wave) % TMC_SERVE_CACHE_COPIES : 0u;

    /* a lane holds a live photon while its weight is >= 0; -1 marks a lane
     * without one (a comparison, so each ballot of it is one vector compare) */
    Photon ph;
    launch_photon(ph, 0, 0u, 0u);
    ph.weight = -1;

    bool more = true;
    while (more) {
        uint64_t dead = __ballot(ph.weight < 0);
        /* conditions are tested one at a time, the common case first: the
         * compiler turned the combined forms into chains of scalar selects
         * evaluated on every step */
        if (dead != 0 && pool_next == pool_end) {
            /* the pool is empty: claim a chunk, or switch sets once drained
             * (PENDING is set by a claim with the pool empty and stays so
             * until the switch) */
            if (SERVE && dead != ~0ull) {
                /* serving: a wave takes its next ticket only once its lanes
                 * have drained.  Reserved earlier, a chunk of another request
                 * (the usual case: most requests are one chunk or a few)
                 * waits out the drain, milliseconds for long walks, while
                 * idle waves could walk it: measured at half load, p50
                 * 10.3 -> 6.0 ms and p99 64 -> 56 ms, for 1% of the capacity
                 * at saturation (a wave whose next chunk would continue its
                 * request now drains first) */
            } else if (!(flags & (PENDING | QUEUE_DONE))) {
                BatchChunk ch;
                const ClaimResult r = try_claim<SERVE>(a, lane, group, ticket, ch);
                if (r == CLAIM_GOT) {
                    fails = 0;
                    if (ch.set == cur_set) {
                        if (SERVE && ch.count == 0) {
                            publish_skip(a, st, lane, ch.first); /* cancelled photons */
                        } else {
                            pool_next = ch.first;
                            pool_end = ch.first + ch.count;
                        }
                    } else {
                        pend = ch;
                        flags |= PENDING;
                    }
                } else if (r == CLAIM_DONE) {
                    flags |= QUEUE_DONE;
                } else if (SERVE) {
                    /* ran dry: publish what was completed, then wait a
                     * little.  A wave never leaves on its own: the ticket it
                     * holds would strand the chunk later published there.
                     * The host asks the launch to stop only once nothing is
                     * active, when every outstanding ticket lies past the
                     * tail. */
                    if (cur_set != 0xFFFFFFFFu) {
                        uint32_t completed = launched + live_pub; /* no lane is live */
                        flush_and_publish<SERVE>(slot, a, st, lane, completed, absorbed, steps);
                        launched = live_pub = 0;
                    }
                    /* back off while dry: 1.7 us after the first failed
                     * poll, doubling to 27 us (a new chunk waits at most
                     * that long for an idle wave) */
                    for (uint32_t k = 0; k < (1u << fails); ++k)
                        __builtin_amdgcn_s_sleep(64);
                    fails = fails < 4u ? fails + 1u : 4u;
                }
            }
            if ((flags & PENDING) && dead == ~0ull) {
                /* drained: hand the finished set's sums over, switch sets */
                if (cur_set != 0xFFFFFFFFu) {
                    uint32_t completed = launched + live_pub; /* no lane is live */
                    flush_and_publish<SERVE>(slot, a, st, lane, completed, absorbed, steps);
                    launched = live_pub = 0;
                }
                cur_set = pend.set;
                load_set<SAMPLER, SERVE>(st, a, cur_set & slot_mask, lane);
                lt.fmax = tally_fmax(st.phys.shell_max, lane & (TMC_SLOT_HOT - 1));
                /* every lane is empty here: only a small cap needs checks */
                flags = st.kcap <= TMC_SLOT_FLUSH_STEPS ? (flags | CAPPING) : (flags & ~CAPPING);
                if (SERVE && pend.count == 0) {
                    publish_skip(a, st, lane, pend.first); /* cancelled photons */
                    pool_next = pool_end = 0;
                } else {
                    pool_next = pend.first;
                    pool_end = pend.first + pend.count;
                }
                flags &= ~PENDING;
            }
        }
        /* the lanes holding a photon: those not dead, and the dead ones that
         * the refill below gives one (all of them unless the pool runs
         * short), kept in scalar registers: the step runs under this mask
         * with no vector compare of its own */
        uint64_t live = ~dead;
        /* the dead lanes take photons from the pool once there are
         * TMC_SLOT_RELAUNCH_MIN of them or the wave has drained */
        const uint32_t n_dead = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)__popcll(dead));
        if (dead != 0 && (n_dead >= TMC_SLOT_RELAUNCH_MIN || dead == ~0ull)) {
            if (pool_next != pool_end) {
                const uint32_t take = pool_take(pool_next, pool_end, n_dead);
                const uint64_t first = st.first_photon + pool_next;
                if (ph.weight < 0 && mask_rank(dead) < take)
                    launch_photon(ph, first + mask_rank(dead), st.phys.key0, st.phys.key1);
                pool_next += take;
                launched += take;
                live = take == n_dead ? ~0ull : __ballot(ph.weight >= 0);
            }
        }
        /* the loop ends with no photon left and no chunk to come (a loop
         * condition rather than a break: with a break, or with this
         * assignment in an else branch of the step, the compiler tested the
         * exit through chains of scalar selects, or copied the photon state
         * around the step) */
        more = live != 0 || (flags & (PENDING | QUEUE_DONE)) != QUEUE_DONE;
        if (live != 0) {
            absorbed += (uint32_t)__popcll(SAMPLER == 1u ? __ballot(ph.weight >= 0 && absorbs<SAMPLER>(ph)) : live);
            ++steps;

            /* the loop has one back-edge (no continue): with several, the
             * compiler copied the photon state in and out of the lane-masked
             * step every iteration */
            if (__builtin_amdgcn_inverse_ballot_w64(live) && !step<SAMPLER, MATH, false, false, true>(ph, st.phys, lt))
                ph.weight = -1;
            /* photons that reached the absorption cap stop; checked only while
             * some photon of the wave can reach it before the next flush */
            if (flags & CAPPING) {
                const uint64_t capped = __ballot(ph.k >= st.kcap) & __ballot(ph.weight >= 0);
                if (capped != 0) {
                    if (lane == 0)
                        atomicAdd(args_at_use<SERVE>(a).counters + 2, (unsigned long long)__popcll(capped));
                    if (ph.k >= st.kcap)
                        ph.weight = -1;
                }
            }

            /* overflow guard: a flush (with a publication) once the
             * absorptions since the last publication (which flushes too)
             * reach TMC_SLOT_FLUSH_DEPOSITS, so no slot word takes more
             * deposits between two flushes.  The cap condition is evaluated
             * at least as often: the other publications happen only when
             * every lane is empty. */
            if (absorbed >= (uint32_t)TMC_SLOT_FLUSH_DEPOSITS) {
                const uint32_t live_now = (uint32_t)__popcll(__ballot(ph.weight >= 0));
                uint32_t completed = launched + live_pub - live_now;
                flush_and_publish<SERVE>(slot, a, st, lane, completed, absorbed, steps);
                launched = 0;
                live_pub = live_now;
                flags = cap_within(ph, ph.weight >= 0, st.kcap, TMC_SLOT_FLUSH_STEPS) ? (flags | CAPPING) : (flags & ~CAPPING);
            }
        }
    }
    if (cur_set != 0xFFFFFFFFu) {
        uint32_t completed = launched + live_pub; /* no lane is live */
        flush_and_publish<SERVE>(slot, a, st, lane, completed, absorbed, steps);
    }
    if (lane == 0) {
        /* serving: the last wave to leave rewinds the ticket counter for the
         * next launch (no wave takes a ticket after it has left) */
        if (SERVE && atomicAdd(a.counters + 3, 1ull) == a.total_waves - 1) {
            atomicExch(a.counters, 0ull);
            atomicExch(a.counters + 3, 0ull);
        }
    }
}

/* Clears the tally pool, the done counters and the batch counters (a
 * kernel rather than a memset so the sequence can be captured in a graph). */
__global__ void batch_clear(unsigned long long* tally, uint64_t tally_words, unsigned long long* done,
                            uint32_t nsets, unsigned long long* counters)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < tally_words; i += stride)
        tally[i] = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * (uint64_t)nsets; i += stride)
        done[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < TMC_BATCH_COUNTERS)
        counters[threadIdx.x] = 0;
}

size_t batch_lds_bytes(uint32_t block, uint32_t shells_max, uint32_t copies)
{
    return (size_t)(block / 64) * slot_words(shells_max, copies) * sizeof(unsigned long long);
}

namespace {

using MultiKernel = void (*)(BatchArgs);

template <bool SERVE>
MultiKernel multi_kernel(uint32_t sampler, uint32_t math)
{
    const bool rej = sampler == 1u;
    if (math == TMC_MATH_OWN)
        return rej ? walk_multi<1u, SERVE, TMC_MATH_OWN> : walk_multi<0u, SERVE, TMC_MATH_OWN>;
    return rej ? walk_multi<1u, SERVE, TMC_MATH_HARDWARE> : walk_multi<0u, SERVE, TMC_MATH_HARDWARE>;
}

} // namespace

hipError_t launch_batch(uint32_t sampler, uint32_t math, const BatchArgs& a, uint32_t grid, uint32_t block,
                        hipStream_t stream)
{
    const size_t lds = batch_lds_bytes(block, a.shells_max, a.copies);
    hipLaunchKernelGGL(multi_kernel<false>(sampler, math), dim3(grid), dim3(block), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_serve(uint32_t sampler, uint32_t math, const BatchArgs& a, uint32_t grid, uint32_t block,
                        hipStream_t stream)
{
    const size_t lds = batch_lds_bytes(block, a.shells_max, a.copies);
    hipLaunchKernelGGL(multi_kernel<true>(sampler, math), dim3(grid), dim3(block), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_batch_clear(unsigned long long* tally, uint64_t tally_words, unsigned long long* done,
                              uint32_t nsets, unsigned long long* counters, hipStream_t stream)
{
    hipLaunchKernelGGL(batch_clear, dim3(1024), dim3(256), 0, stream, tally, tally_words, done, nsets, counters);
    return hipGetLastError();
}

hipError_t batch_occupancy(uint32_t sampler, uint32_t math, uint32_t block, uint32_t shells_max, uint32_t copies,
                           int* wg_per_cu)
{
    const size_t lds = batch_lds_bytes(block, shells_max, copies);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, multi_kernel<false>(sampler, math), (int)block,
                                                        lds);
}

hipError_t serve_occupancy(uint32_t sampler, uint32_t math, uint32_t block, uint32_t shells_max, uint32_t copies,
                           int* wg_per_cu)
{
    const size_t lds = batch_lds_bytes(block, shells_max, copies);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, multi_kernel<true>(sampler, math), (int)block,
                                                        lds);
}

} // namespace tmc
