// Host-visible launch interface of the HIP kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"

namespace mdft {

// One axis pass of one image inside a batch.
struct BatchJob {
    AxisJob job;
    int logm;
    int blue;
    int cols;
    int ntiles;
    int dep;       // counter to wait for (-1: none)
    int dep_need;  // value the counter must reach
    int sig;       // counter to bump after each tile (-1: none)
    int pad;
};

// ctl layout: [0] ticket, [1] error flag, [2..3] reserved, then counters
constexpr uint32_t BATCH_CTL_HEADER = 4;
// item.y flag: a whole small image, item.x = row job, item.y & ~flag = column job
constexpr uint32_t BATCH_WHOLE = 0x80000000u;

hipError_t gen_twiddles(float2* table, hipStream_t stream);
hipError_t gen_chirp(double2* chirp64, float2* chirp32, int n, hipStream_t stream);
hipError_t gen_dtab(float2* dtab, const double2* chirp64, double* cos_scratch, int n, int m, hipStream_t stream);

hipError_t launch_axis(int logm, const AxisJob& job, bool blue, bool cols, bool wide, hipStream_t stream);

// batch kernels: threads is 256 (padded lengths up to 4096) or 1024 (up to 8192)
hipError_t launch_batch_grid(int threads, const BatchJob* jobs, const uint2* items, uint32_t nitems,
                             hipStream_t stream);
hipError_t launch_batch_persistent(int threads, const BatchJob* jobs, const uint2* items, uint32_t nitems,
                                   uint32_t* ctl, uint32_t nctl, int grid, uint32_t spin_limit, hipStream_t stream);
int batch_lines_per_tile(int threads, int logm);
hipError_t batch_occupancy(int threads, int* blocks_per_cu);
int batch_max_logm(int threads);

}  // namespace mdft
