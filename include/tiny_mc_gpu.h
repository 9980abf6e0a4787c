/* tiny_mc_gpu — C ABI of the MI355X photon-transport path.
 *
 * Drop-in for the reference's photon transport loop (computacionparalela/
 * tiny_mc): where tiny_mc.c:43-49 seeds glibc rand() and calls
 *     void photon(float* heats, float* heats_squared);      (photon.h:3)
 * once per photon, a host program fills a tmc_params from the same params.h
 * macros and makes one call that walks every photon on the GPU.
 *
 * Conventions
 *  - Plain C types only; callers own every array.
 *  - Functions return TMC_OK (0) or a negative tmc_status; none aborts or
 *    prints.  tiny_mc_gpu_strerror() names a status.
 *  - One implicit context per process on the HIP device current at the first
 *    call (or the one given to tiny_mc_gpu_init); calls are serialised
 *    internally and synchronous.
 *  - Results are deterministic: the random stream is Philox4x32-10 keyed by
 *    (seed, photon index) and the tallies are integers, so a run gives the
 *    same bits on every launch shape and in every split of the photon range.
 */
#ifndef TINY_MC_GPU_H
#define TINY_MC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TMC_ABI_VERSION 1

/* Largest shell count (LDS holds 16 bytes per shell per workgroup). */
#define TMC_MAX_SHELLS 8192

/* Fixed-point scales of the raw tallies (see tiny_mc_gpu_run_ex). */
#define TMC_HEAT_FRAC_BITS 33
#define TMC_HEAT2_FRAC_BITS 43

typedef enum tmc_status {
    TMC_OK = 0,
    TMC_E_INVALID = -1, /* bad parameters or null pointer */
    TMC_E_NODEVICE = -2, /* no HIP device usable */
    TMC_E_HIP = -3, /* a HIP runtime call failed */
    TMC_E_NOMEM = -4, /* device or host allocation failed */
    TMC_E_RANGE = -5, /* photon index range first_photon + photons exceeds 2^63 */
    TMC_E_UNSUPPORTED = -6, /* option or library build not available */
    TMC_E_TOO_LARGE = -7, /* more shells than the static scratch of tiny_mc_gpu_run_noalloc holds */
    TMC_E_CANCELLED = -8, /* serving: the request was cancelled (tmc_serve_cancel, tmc_serve_close_timeout) */
    TMC_E_DEADLINE = -9, /* serving: the request's deadline passed before it finished */
    TMC_E_TIMEOUT = -10 /* serving: tmc_serve_close_timeout ran out of time; the server is still open */
} tmc_status;

/* Physics and size of one simulation; the fields are the macros of
 * params.h:5-27 (SEED is the separate seed argument). */
typedef struct tmc_params {
    float mu_a; /* MU_A: absorption coefficient, 1/cm, > 0 */
    float mu_s; /* MU_S: scattering coefficient, 1/cm, >= 0 */
    uint32_t shells; /* SHELLS: 1..TMC_MAX_SHELLS; the last shell collects
                               every deposit beyond the others (tiny_mc.c:66) */
    float microns_per_shell; /* MICRONS_PER_SHELL: > 0 */
    uint64_t photons; /* PHOTONS */
} tmc_params;

/* Direction sampler after each absorption (photon.c:34-43). */
#define TMC_SAMPLER_ANALYTIC 0u /* cos(theta) = 2u-1, phi = 2 pi u' (default) */
#define TMC_SAMPLER_REJECTION 1u /* the reference's rejection method */

/* Arithmetic of the walk (the tallies are the same integers in every mode).
 * The reference computes in binary32 (photon.c); the other modes exist to
 * measure what that costs where it matters most (albedo near 1). */
#define TMC_PRECISION_FP32 0u /* the reference's binary32 arithmetic (default) */
#define TMC_PRECISION_FP32_WEIGHT 1u /* binary32, weight from exact counts and FP64 constants */
#define TMC_PRECISION_FP32_COMPENSATED 2u /* as _WEIGHT, plus compensated position sums */
#define TMC_PRECISION_FP64 3u /* every quantity in binary64 */

/* How the binary32 walk evaluates its two transcendental quantities, the
 * step length -ln u (photon.c:21) and the direction's sin/cos of 2 pi u'.
 * Both are accurate to about one unit in the last place (the hardware log:
 * <= 1.92 ulp on every input the walk can produce; sin/cos: <= 1.3e-7
 * absolute); the own polynomials exist so that a host program can
 * reproduce the walk bit for bit without evaluating anything on the device. */
#define TMC_MATH_HARDWARE 0u /* v_log_f32, v_sin_f32, v_cos_f32 (default) */
#define TMC_MATH_OWN 1u /* own minimax polynomials (tmc_math.h) */

typedef struct tmc_options {
    uint32_t sampler; /* TMC_SAMPLER_* */
    uint32_t block; /* threads per workgroup, multiple of 64 in [64, 1024]; 0 = auto */
    uint32_t grid; /* workgroups; 0 = auto (fills every CU) */
    uint32_t chunk; /* photons a wave claims at a time, multiple of 64; 0 = auto */
    uint64_t first_photon; /* index of the first photon of the stream to walk */
    uint64_t max_interactions; /* stop a photon after this many absorptions; 0 = never */
    uint32_t precision; /* TMC_PRECISION_* */
    uint32_t math; /* TMC_MATH_* (binary32 walk of TMC_PRECISION_FP32; the other
                      precision modes use the own polynomials or binary64) */
    uint32_t weight_table; /* binary32 walk: the weight of a surviving photon is a function of
                              its absorption count, so the walk can read its fixed-point images
                              from a per-workgroup table instead of computing them (same bits).
                              0 = automatic (runs of 2^20 photons or more, either sampler, whose
                              table spans the first roulette cycle and eight more, up to 1024
                              entries);
                              TMC_WEIGHT_TABLE_OFF = never; N = always, at most N entries */
    uint32_t reserved; /* 0 */
} tmc_options;

#define TMC_WEIGHT_TABLE_OFF 0xFFFFFFFFu

typedef struct tmc_run_stats {
    uint64_t photons; /* photons walked */
    uint64_t interactions; /* absorption events (iterations of photon.c:20) */
    uint64_t truncated; /* photons stopped by max_interactions */
    double kernel_ms; /* walk kernel time, HIP events on the library stream */
    double total_ms; /* whole call: clears, kernel, copies back */
    uint32_t grid, block, chunk; /* launch shape used */
    uint32_t sampler;
    uint32_t precision;
    uint32_t math;
    uint32_t weight_table; /* entries of the weight table used (0: none) */
} tmc_run_stats;

/* Default options (analytic sampler, automatic launch shape, photon 0). */
void tmc_options_default(tmc_options* opt);

/* Walks p->photons photons of stream `seed` and writes, for every shell s,
 *   heat[s]  = sum over deposits in s of (1 - albedo) * w
 *   heat2[s] = sum over deposits in s of (1 - albedo)^2 * w^2
 * i.e. the reference's heats / heats_squared (photon.c:30-31) after the
 * whole loop (tiny_mc.c:47-49), in double.  Arrays have p->shells elements
 * and are overwritten.  Replaces srand(SEED) + the photon() loop. */
int tiny_mc_gpu_run(const tmc_params* p, uint64_t seed, double* heat, double* heat2);

/* As tiny_mc_gpu_run with options, returning the exact integer tallies:
 * tally has 4 * p->shells words; words [2s, 2s+1] are the (low, high) halves
 * of the 128-bit sum of w * 2^TMC_HEAT_FRAC_BITS over deposits in shell s and
 * words [2S + 2s, 2S + 2s + 1] those of (w*w) * 2^TMC_HEAT2_FRAC_BITS
 * (w in binary32: both images are exact for w >= 2^-10, which every deposit
 * is at albedos of 0.1 or more; below that, deposits of photons that
 * survived several roulettes can be smaller and their images are truncated,
 * deterministically; TMC_PRECISION_FP64: rounded to nearest).  Tallies of
 * disjoint photon ranges of one stream add as integers.  stats may be
 * NULL. */
int tiny_mc_gpu_run_ex(const tmc_params* p, uint64_t seed, const tmc_options* opt, uint64_t* tally,
                       tmc_run_stats* stats);

/* Largest shell count of tiny_mc_gpu_run_noalloc: its static device scratch
 * holds 4 + 4 * TMC_NOALLOC_MAX_SHELLS 64-bit words (128 KiB) per device. */
#define TMC_NOALLOC_MAX_SHELLS 4096

/* tiny_mc_gpu_run_ex without any device or host allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate (code that runs
 * while a HIP graph is captured elsewhere in the process, tight serving
 * loops, hosts that forbid allocation after start-up).  The walk's counters
 * and tallies live in a static device array of fixed size, one copy per
 * device, held until the process ends; its address is looked up once per
 * device when the context is created (tiny_mc_gpu_init or the first call).
 * Results are bit-identical to tiny_mc_gpu_run_ex with the same arguments.
 * Returns TMC_E_TOO_LARGE for p->shells > TMC_NOALLOC_MAX_SHELLS.  Every
 * call on a device uses that device's one array, which keeps the last
 * call's contents (each call clears what it uses first); the library
 * serialises its calls, so calls from several threads of a process wait for
 * each other, and code that bypassed that lock would race on the array. */
int tiny_mc_gpu_run_noalloc(const tmc_params* p, uint64_t seed, const tmc_options* opt, uint64_t* tally,
                            tmc_run_stats* stats);

/* Converts raw tallies (as from tiny_mc_gpu_run_ex) into heat / heat2 in
 * the reference's units, in double (host only, no device work). */
int tiny_mc_gpu_tally_to_heat(const tmc_params* p, const uint64_t* tally, double* heat, double* heat2);

/* Selects the HIP device for the implicit context (optional; default: the
 * device current at the first call).  Fails if a context already exists on
 * another device. */
int tiny_mc_gpu_init(int device);

/* Releases the implicit context and its device memory. */
void tiny_mc_gpu_shutdown(void);

/* Device name (gcnArchName) and compute-unit count of the context's device. */
int tiny_mc_gpu_device_info(char* name, size_t cap, int* compute_units);

/* Expected absorptions per photon for p (the binary32 weight cadence of
 * photon.c:32,45-49 and the roulette survival probability): the estimate the
 * batched and served modes schedule with.  Host only. */
double tiny_mc_gpu_expected_interactions(const tmc_params* p);

/* Human-readable name of a tmc_status. */
const char* tiny_mc_gpu_strerror(int status);

/* Diagnostics used by the parity tests: the device Philox4x32-10 on n
 * (counter, key) pairs (ctr: 4n words, key: 2n words, out: 4n words), and the
 * device log / sin(2 pi a) / cos(2 pi a) of the walk on n inputs. */
int tiny_mc_gpu_philox(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out);
int tiny_mc_gpu_eval_math(const float* in, uint32_t n, float* log_out, float* sin_out, float* cos_out);

/* The walk's own step length -ln(u), sin/cos(2 pi u) and the analytic
 * sampler's sin(theta) = sqrt(1 - (2u - 1)^2) in math mode `math`
 * (TMC_MATH_*) on n inputs: the exact values a binary32 walk uses. */
int tiny_mc_gpu_eval_step_math(uint32_t math, const float* in, uint32_t n, float* step_out, float* sin_out,
                               float* cos_out, float* polar_out);

/* Diagnostic: compares the walk's short correctly rounded square root with
 * the device's IEEE square root on every float whose bit pattern lies in
 * [first_bits, first_bits + count); returns the number of mismatches and
 * the lowest mismatching pattern (0 if none). */
int tiny_mc_gpu_check_sqrt(uint32_t first_bits, uint64_t count, uint64_t* mismatches, uint32_t* first_mismatch);

/* Diagnostic: the same comparison for each short form the walk uses in
 * place of an IEEE operation:
 *   TMC_CHECK_SQRT             the short square root (= tiny_mc_gpu_check_sqrt);
 *   TMC_CHECK_DIV_TENTH_SHORT  fma(w, 10, w * C_lo) against w / 0.1f (the
 *                              roulette's weight update), exact for w >= 2^-100;
 *   TMC_CHECK_DIV_TENTH        the guarded form the walk calls, exact for all w. */
#define TMC_CHECK_SQRT 0u
#define TMC_CHECK_DIV_TENTH_SHORT 1u
#define TMC_CHECK_DIV_TENTH 2u
int tiny_mc_gpu_check_short_op(uint32_t op, uint32_t first_bits, uint64_t count, uint64_t* mismatches,
                               uint32_t* first_mismatch);

/* Diagnostic: the binary32 walk takes a deposit's shell from the hardware
 * square root where that is certain to give the shell of the correctly
 * rounded one (photon.c:26-29), its coordinate clamped to a lane's bound
 * past the last shell, which picks that lane's copy of the overflow shell.
 * Checks that claim for the physics p on every squared radius whose bit
 * pattern lies in [first_bits, first_bits + count) and for each of the 32
 * copies: violations (radii marked certain whose inner shell differs, or
 * that land on the wrong side of the overflow boundary or outside the
 * copies) with the lowest such pattern, and the number of radii not marked
 * certain for copy 0 (the walk recomputes those with the IEEE square
 * root). */
int tiny_mc_gpu_check_shell_index(const tmc_params* p, uint32_t first_bits, uint64_t count, uint64_t* violations,
                                  uint32_t* first_violation, uint64_t* unsure);

/* ------------------------------------------------------------------------
 * Batched sweep: many parameter sets in one persistent launch (one per
 * LDS size class of the sets, up to three, back to back).
 *
 * No counterpart in the reference (it has one compile-time parameter set);
 * this is the loop of tiny_mc.c:47-49 run for many params.h configurations
 * at once.  Each set's tallies are bit-identical to tiny_mc_gpu_run_ex on
 * that set alone (binary32 arithmetic, TMC_PRECISION_FP32).
 * ------------------------------------------------------------------------ */

/* One simulation of a batch: physics and photon count, its seed, and the
 * stream index of its first photon. */
typedef struct tmc_set {
    tmc_params params;
    uint64_t seed;
    uint64_t first_photon;
} tmc_set;

typedef struct tmc_batch_options {
    uint32_t sampler; /* TMC_SAMPLER_* for every set */
    uint32_t block; /* threads per workgroup, multiple of 64; 0 = auto (256) */
    uint32_t waves_per_simd; /* resident waves per SIMD, 1..8; 0 = as many as fit */
    uint32_t copies; /* LDS tally copies per wave: 1 or 2; 0 = auto (1) */
    uint64_t chunk_work; /* target absorptions per queue chunk; 0 = auto */
    uint64_t max_interactions; /* per photon, 0 = never stop early */
    uint32_t use_graph; /* nonzero: plans replay a captured HIP graph */
    uint32_t math; /* TMC_MATH_* */
} tmc_batch_options;

typedef struct tmc_batch_stats {
    uint64_t sets;
    uint64_t photons;
    uint64_t interactions;
    uint64_t truncated;
    uint64_t chunks; /* queue length */
    uint64_t chunk_work; /* target absorptions per chunk used */
    double expected_interactions; /* host estimate used for scheduling */
    double kernel_ms; /* clear + walk on the device (HIP events) */
    double total_ms; /* whole call */
    uint32_t grid, block, copies, waves_per_simd;
    uint64_t idle_lane_steps; /* lanes without a photon while their wave stepped
                                 (drains at set switches and the end of the queue):
                                 lane utilisation = interactions / (interactions + idle) */
    uint32_t passes; /* launches, one per LDS size class of the sets (grid, block,
                        waves_per_simd, chunk_work above: the pass with the most work) */
    uint32_t reserved;
} tmc_batch_stats;

typedef struct tmc_batch_plan tmc_batch_plan;

void tmc_batch_options_default(tmc_batch_options* opt);

/* Words of the concatenated tallies of a batch: 4 * sum of shells.  Set i's
 * tally (layout as tiny_mc_gpu_run_ex) starts at word 4 * sum_{j<i} shells_j. */
uint64_t tmc_batch_tally_words(const tmc_set* sets, uint32_t nsets);

/* Validates the sets, builds the chunk queues (longest first) and uploads the
 * set table and queues to the device; with use_graph, captures the steady
 * launch sequence (clear, walks, copy back) as a HIP graph.  Sets whose
 * shell counts allow more resident waves (smaller per-wave LDS slots) run in
 * a launch of their own when they carry at least a tenth of the work. */
int tmc_batch_plan_create(const tmc_set* sets, uint32_t nsets, const tmc_batch_options* opt,
                          tmc_batch_plan** plan);

/* Runs the plan; tallies receives tmc_batch_tally_words() words. */
int tmc_batch_plan_run(tmc_batch_plan* plan, uint64_t* tallies, tmc_batch_stats* stats);

void tmc_batch_plan_destroy(tmc_batch_plan* plan);

/* create + run + destroy. */
int tmc_batch_run(const tmc_set* sets, uint32_t nsets, const tmc_batch_options* opt, uint64_t* tallies,
                  tmc_batch_stats* stats);

/* Picks block size, resident waves, tally copies and chunk work for this
 * batch by timing candidate launches on a reduced copy of it (every set's
 * photon budget divided by `sample_divisor`, at least 64 photons; each
 * candidate the best of three runs).  Starting from *base, a candidate
 * replaces the current choice only when it is at least 3% faster, and a
 * final choice other than *base must be 3% faster than *base again on a
 * copy four times larger (the whole batch when sample_divisor <= 4), so the
 * result stays *base unless both show a clear gain.  Writes the choice to
 * *best. */
int tmc_batch_autotune(const tmc_set* sets, uint32_t nsets, const tmc_batch_options* base,
                       uint32_t sample_divisor, tmc_batch_options* best);

/* ------------------------------------------------------------------------
 * Serving: a request loop over a resident persistent kernel.
 *
 * No counterpart in the reference.  Requests (one tmc_set each) are admitted
 * in arrival order while a request slot and room in the tally arena are
 * free; their chunks are published, shortest remaining work first, into a
 * ring in host-coherent memory that the running kernel polls, so requests
 * that arrive while others are on the GPU join the same launch.  Every
 * result is bit-identical to tiny_mc_gpu_run_ex of the same set alone
 * (binary32 arithmetic, TMC_PRECISION_FP32).  A server owns a host thread
 * and one default-priority stream on the context's device; while it holds
 * requests its kernel occupies the device's workgroup slots but
 * reserve_per_cu per compute unit (and a few more).  Other kernels on the
 * device (single runs, batches, another server's) get the workgroups of
 * their launch spread over every compute unit, so with the default of none
 * reserved they wait until the server has been idle for idle_ms and its
 * kernel has left; with 2 per compute unit they run beside it.
 * ------------------------------------------------------------------------ */

#define TMC_PENDING 1 /* positive status: the request has not finished */

typedef struct tmc_serve_options {
    uint32_t sampler; /* TMC_SAMPLER_* for every request */
    uint32_t max_shells; /* largest shell count of a request; 0 = 401 */
    uint32_t max_requests; /* resident requests (slots), <= 65536; 0 = 4096 */
    uint32_t block; /* threads per workgroup; 0 = 256 */
    uint64_t tally_words; /* device tally arena in 64-bit words; 0 = max_requests * 4 * max_shells */
    uint64_t ring_chunks; /* chunk ring entries, power of two >= 2; 0 = 2^20 */
    uint64_t chunk_work; /* target absorptions per chunk at light load; 0 = 2^17 */
    uint32_t max_photons_per_lane; /* cap on photons per lane in a chunk at light load, <= 2^24; 0 = 16 */
    uint32_t idle_ms; /* the server stops the kernel after this long without requests; 0 = 50 */
    uint32_t math; /* TMC_MATH_* */
    uint32_t max_chunk_scale; /* under load (unpublished work beyond two chunks per resident wave)
                                 chunk_work and max_photons_per_lane grow with the backlog, up to
                                 this factor (fewer wave drains per absorption); one request's
                                 work counts up to four chunks per wave, so a request alone grows
                                 its chunks at most 2x; 1 = fixed chunks, 0 = 128, <= 1024 */
    uint32_t reserve_per_cu; /* workgroup slots per compute unit the resident kernel leaves free
                                for other work on the device; 0 = none (the whole device
                                serves); 2 lets single runs and batches run beside the server */
    uint32_t reserved; /* 0 */
} tmc_serve_options;

typedef struct tmc_request_stats {
    uint64_t photons;
    uint64_t interactions;
    double submit_s; /* seconds since the server opened */
    double admit_s;
    double done_s;
    double latency_ms; /* done - submit */
} tmc_request_stats;

typedef struct tmc_serve_stats {
    uint64_t submitted, admitted, completed;
    uint64_t photons_completed, interactions_completed;
    uint64_t kernel_launches;
    uint64_t chunks_published;
    uint64_t chunks_claimed; /* ring indices handed to waves (last observed) */
    uint64_t pending, active; /* requests waiting for admission / admitted, unfinished */
    double elapsed_s;
    uint64_t cancelled; /* requests cancelled by tmc_serve_cancel or a bounded close */
    uint64_t expired; /* requests cancelled by their deadline */
    uint32_t grid, block; /* launch shape of the resident kernel */
} tmc_serve_stats;

typedef struct tmc_server tmc_server;

void tmc_serve_options_default(tmc_serve_options* opt);

int tmc_serve_open(const tmc_serve_options* opt, tmc_server** server);

/* Queues a request; returns its id.  TMC_E_INVALID for bad parameters or
 * more shells than max_shells, TMC_E_NOMEM if its tally can never fit. */
int tmc_serve_submit(tmc_server* server, const tmc_set* request, uint64_t* id);

/* As tmc_serve_submit, with a deadline: a request not finished deadline_s
 * seconds after its submission is cancelled (poll / wait then return
 * TMC_E_DEADLINE).  deadline_s <= 0: none. */
int tmc_serve_submit_ex(tmc_server* server, const tmc_set* request, double deadline_s, uint64_t* id);

/* Cancels a request whose outcome has not been handed over: poll / wait
 * then return TMC_E_CANCELLED (once).  A queued request is dropped; an
 * admitted one publishes no further chunks, and its slot and arena region
 * are reused once the chunks already published are done (every other
 * request's result is unaffected); a finished one loses its result.
 * TMC_E_INVALID for an unknown id. */
int tmc_serve_cancel(tmc_server* server, uint64_t id);

/* TMC_OK and the request's tally (4 * shells words, layout of
 * tiny_mc_gpu_run_ex) once it is done — the outcome is handed over once —,
 * TMC_PENDING before, TMC_E_CANCELLED / TMC_E_DEADLINE for a cancelled
 * request, TMC_E_INVALID for an unknown id. */
int tmc_serve_poll(tmc_server* server, uint64_t id, uint64_t* tally, tmc_request_stats* stats);

/* As tmc_serve_poll, blocking up to timeout_s seconds (< 0: no limit). */
int tmc_serve_wait(tmc_server* server, uint64_t id, double timeout_s, uint64_t* tally, tmc_request_stats* stats);

int tmc_serve_get_stats(tmc_server* server, tmc_serve_stats* stats);

/* Finishes every submitted request, stops the kernel, frees the server.
 * Results not yet polled are discarded. */
int tmc_serve_close(tmc_server* server);

/* tmc_serve_close within a time limit: requests still unfinished timeout_s
 * seconds after the call are cancelled; the chunks already claimed drain
 * and the kernel leaves.  If that has not happened within another timeout_s
 * (plus one second), returns TMC_E_TIMEOUT with the server still open (it
 * accepts no new requests; call again to keep waiting).  timeout_s < 0: no
 * limit (= tmc_serve_close). */
int tmc_serve_close_timeout(tmc_server* server, double timeout_s);

#ifdef __cplusplus
}
#endif

#endif
