/* Internal interface between the C ABI (tmc_api.cpp) and the kernels
 * (tmc_walk.hip).  Not installed; the public interface is
 * include/tiny_mc_gpu.h. */
#ifndef TMC_INTERNAL_H
#define TMC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tmc_plan.h"

namespace tmc {

/* Everything one launch of the walk kernel needs, passed by value. */
struct WalkArgs {
    uint64_t seed; /* Philox key */
    uint64_t first_photon; /* photon index of claim offset 0 */
    uint64_t photons; /* photons in this launch */
    float albedo; /* mu_s / (mu_s + mu_a) in binary32, as photon.c:8 */
    float spm; /* shells per mean free path, as photon.c:9 */
    float shell_max; /* (float)(shells - 1) */
    uint32_t shells;
    uint32_t chunk; /* photons claimed per wave at a time (multiple of 64) */
    uint32_t flush_interval; /* wave steps between LDS tally flushes */
    uint32_t kcap; /* stop a photon after kcap absorptions */
    uint32_t table_k; /* entries of the weight table built in LDS (binary32 walk); 0: none */
    unsigned long long* counters; /* [0] claim offset, [1] interactions, [2] truncated */
    unsigned long long* tally; /* 4*shells words of 128-bit sums (lo, hi) */
    /* precision variants */
    double albedo_d; /* mu_s / (mu_s + mu_a) in FP64 */
    double spm_d; /* 1e4 / microns_per_shell / (mu_a + mu_s) in FP64 */
    float l2a_hi, l2a_lo; /* log2(albedo_d) as a float pair */
    float l2s_hi, l2s_lo; /* -log2((double)0.1f) as a float pair */
    float spm_w; /* spm_d rounded to float */
    uint32_t pad2;
};

/* Launches the persistent walk kernel of `precision` (TMC_PRECISION_*) on
 * `stream`.  LDS: 16 bytes per shell plus the overflow shell's lane copies,
 * and a.table_k entries of the weight table (TMC_PRECISION_FP32 only). */
hipError_t launch_walk(uint32_t sampler, uint32_t precision, uint32_t math, const WalkArgs& a, uint32_t grid,
                       uint32_t block, hipStream_t stream);

/* Resident workgroups per CU for the walk kernel at this block size. */
hipError_t walk_occupancy(uint32_t sampler, uint32_t precision, uint32_t math, uint32_t block, uint32_t shells,
                          uint32_t table_k, int* wg_per_cu);

/* Zeroes n 64-bit words at p on `stream` (a kernel, so the sequence holds no
 * memset node). */
hipError_t launch_clear(unsigned long long* p, size_t n, hipStream_t stream);

hipError_t launch_philox(const uint32_t* ctr, const uint32_t* key, uint32_t n, uint32_t* out, hipStream_t stream);
hipError_t launch_check_shell(float spm, float shell_max, uint32_t first, uint64_t count, unsigned long long* out,
                              hipStream_t stream);
hipError_t launch_check_short_op(uint32_t op, uint32_t first, uint64_t count, unsigned long long* out,
                                 hipStream_t stream);
hipError_t launch_eval_step_math(uint32_t math, const float* in, uint32_t n, float* t, float* sn, float* cs,
                                 float* sp, hipStream_t stream);
hipError_t launch_eval_math(const float* in, uint32_t n, float* lg, float* sn, float* cs, hipStream_t stream);

} // namespace tmc

#endif
