// Axis-pass kernels: one launch transforms every line of one axis of one
// image (the reference records rows_to_buffer, FFT, mul_d, FFT,
// rows_from_buffer as five dispatches per axis, src/gpu/dft/dft.cpp:326-347;
// here they are one fused kernel per axis), and the persistent batch kernel
// that runs the axis passes of many images of mixed sizes in one launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"
#include "launch.hpp"

namespace mdft {

// Workgroup size of the single-image kernels: 256 threads, or one line when a
// line needs more threads (M = 8192 -> 512, M = 16384 -> 1024).
__host__ __device__ constexpr int axis_threads(int logm)
{
    return (1 << logm) / ept_of(logm) > 256 ? (1 << logm) / ept_of(logm) : 256;
}

template <int LOGM, bool BLUE, bool COLS>
__global__ __launch_bounds__(axis_threads(LOGM)) void k_axis(AxisJob job)
{
    constexpr int T = axis_threads(LOGM);
    constexpr int LF = LineCfg<LOGM, T>::template lds_floats<COLS>();
    __shared__ float lds[LF > 0 ? LF : 1];
    const int tile = COLS ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    line_tile<LOGM, T, BLUE, COLS>(job, tile, lds, threadIdx.x);
}

template <int LOGM>
static hipError_t launch_axis_t(const AxisJob& job, bool blue, bool cols, hipStream_t stream)
{
    constexpr int T = axis_threads(LOGM);
    constexpr int LPW = LineCfg<LOGM, T>::LPW;
    const int grid = (job.count + LPW - 1) / LPW;
    if (grid <= 0) return hipSuccess;
    if (blue) {
        if constexpr (LOGM >= 2) {
            if (cols)
                hipLaunchKernelGGL((k_axis<LOGM, true, true>), dim3(grid), dim3(T), 0, stream, job);
            else
                hipLaunchKernelGGL((k_axis<LOGM, true, false>), dim3(grid), dim3(T), 0, stream, job);
        } else {
            return hipErrorInvalidValue;
        }
    } else {
        if (cols)
            hipLaunchKernelGGL((k_axis<LOGM, false, true>), dim3(grid), dim3(T), 0, stream, job);
        else
            hipLaunchKernelGGL((k_axis<LOGM, false, false>), dim3(grid), dim3(T), 0, stream, job);
    }
    return hipGetLastError();
}

hipError_t launch_axis(int logm, const AxisJob& job, bool blue, bool cols, hipStream_t stream)
{
    switch (logm) {
    case 1: return launch_axis_t<1>(job, blue, cols, stream);
    case 2: return launch_axis_t<2>(job, blue, cols, stream);
    case 3: return launch_axis_t<3>(job, blue, cols, stream);
    case 4: return launch_axis_t<4>(job, blue, cols, stream);
    case 5: return launch_axis_t<5>(job, blue, cols, stream);
    case 6: return launch_axis_t<6>(job, blue, cols, stream);
    case 7: return launch_axis_t<7>(job, blue, cols, stream);
    case 8: return launch_axis_t<8>(job, blue, cols, stream);
    case 9: return launch_axis_t<9>(job, blue, cols, stream);
    case 10: return launch_axis_t<10>(job, blue, cols, stream);
    case 11: return launch_axis_t<11>(job, blue, cols, stream);
    case 12: return launch_axis_t<12>(job, blue, cols, stream);
    case 13: return launch_axis_t<13>(job, blue, cols, stream);
    case 14: return launch_axis_t<14>(job, blue, cols, stream);
    default: return hipErrorInvalidValue;
    }
}

int axis_lines_per_tile(int logm)
{
    switch (logm) {
#define MDFT_LPW_CASE(L) \
    case L: return LineCfg<L, axis_threads(L)>::LPW;
        MDFT_LPW_CASE(1)
        MDFT_LPW_CASE(2)
        MDFT_LPW_CASE(3)
        MDFT_LPW_CASE(4)
        MDFT_LPW_CASE(5)
        MDFT_LPW_CASE(6)
        MDFT_LPW_CASE(7)
        MDFT_LPW_CASE(8)
        MDFT_LPW_CASE(9)
        MDFT_LPW_CASE(10)
        MDFT_LPW_CASE(11)
        MDFT_LPW_CASE(12)
        MDFT_LPW_CASE(13)
        MDFT_LPW_CASE(14)
#undef MDFT_LPW_CASE
    default: return 0;
    }
}

// ---------------------------------------------------------------------------
// persistent batch kernel
// ---------------------------------------------------------------------------
// Work items are (job, tile) pairs in a host-chosen order; workgroups take
// the next ticket from a device counter until the list is exhausted.  A
// column job waits for its image's row job: row tiles bump a per-image
// counter after a device-scope release fence; a column tile spins (bounded)
// until the counter reaches the row job's tile count, then acquires.  Every
// item waits only on items with smaller tickets, which were taken by running
// workgroups that never wait on later ones, so the queue always drains.

constexpr int BATCH_THREADS = 512;  // M <= 8192 in a batch

template <int LOGM>
constexpr int batch_lds_floats()
{
    constexpr int a = LineCfg<LOGM, BATCH_THREADS>::template lds_floats<false>();
    constexpr int b = LineCfg<LOGM, BATCH_THREADS>::template lds_floats<true>();
    return a > b ? a : b;
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int BATCH_LDS = cmax(
    cmax(cmax(batch_lds_floats<5>(), batch_lds_floats<6>()), cmax(batch_lds_floats<7>(), batch_lds_floats<8>())),
    cmax(cmax(batch_lds_floats<9>(), batch_lds_floats<10>()),
         cmax(cmax(batch_lds_floats<11>(), batch_lds_floats<12>()), batch_lds_floats<13>())));

template <int LOGM>
__device__ __forceinline__ void batch_run(const BatchJob& bj, int tile, float* lds)
{
    if constexpr (LOGM >= 2) {
        if (bj.blue) {
            if (bj.cols)
                line_tile<LOGM, BATCH_THREADS, true, true>(bj.job, tile, lds, threadIdx.x);
            else
                line_tile<LOGM, BATCH_THREADS, true, false>(bj.job, tile, lds, threadIdx.x);
            return;
        }
    }
    if (bj.cols)
        line_tile<LOGM, BATCH_THREADS, false, true>(bj.job, tile, lds, threadIdx.x);
    else
        line_tile<LOGM, BATCH_THREADS, false, false>(bj.job, tile, lds, threadIdx.x);
}

__global__ __launch_bounds__(BATCH_THREADS) void k_batch(const BatchJob* __restrict__ jobs,
                                                         const uint2* __restrict__ items, uint32_t nitems,
                                                         uint32_t* ctl, uint32_t spin_limit)
{
    __shared__ float lds[BATCH_LDS];
    __shared__ uint32_t s_item;
    uint32_t* ticket = ctl;
    uint32_t* err = ctl + 1;
    uint32_t* counters = ctl + BATCH_CTL_HEADER;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const uint2 item = items[it];
        const BatchJob bj = jobs[item.x];
        if (bj.dep >= 0) {
            if (threadIdx.x == 0) {
                uint32_t spins = 0;
                while (__hip_atomic_load(&counters[bj.dep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       (uint32_t)bj.dep_need) {
                    __builtin_amdgcn_s_sleep(2);
                    if (++spins > spin_limit) {
                        atomicOr(err, 1u);
                        break;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // see the row tiles' stores
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        switch (bj.logm) {
        case 1: batch_run<1>(bj, item.y, lds); break;
        case 2: batch_run<2>(bj, item.y, lds); break;
        case 3: batch_run<3>(bj, item.y, lds); break;
        case 4: batch_run<4>(bj, item.y, lds); break;
        case 5: batch_run<5>(bj, item.y, lds); break;
        case 6: batch_run<6>(bj, item.y, lds); break;
        case 7: batch_run<7>(bj, item.y, lds); break;
        case 8: batch_run<8>(bj, item.y, lds); break;
        case 9: batch_run<9>(bj, item.y, lds); break;
        case 10: batch_run<10>(bj, item.y, lds); break;
        case 11: batch_run<11>(bj, item.y, lds); break;
        case 12: batch_run<12>(bj, item.y, lds); break;
        case 13: batch_run<13>(bj, item.y, lds); break;
        default: break;
        }
        if (bj.sig >= 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // release this wave's stores
            __syncthreads();
            if (threadIdx.x == 0) atomicAdd(&counters[bj.sig], 1u);
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t value)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_batch(const BatchJob* jobs, const uint2* items, uint32_t nitems, uint32_t* ctl, uint32_t nctl,
                        int grid, uint32_t spin_limit, hipStream_t stream)
{
    // clear ticket, error flag and counters with a kernel (graph-capture safe)
    hipLaunchKernelGGL(k_fill_u32, dim3((nctl + 255) / 256), dim3(256), 0, stream, ctl, nctl, 0u);
    hipLaunchKernelGGL(k_batch, dim3(grid), dim3(BATCH_THREADS), 0, stream, jobs, items, nitems, ctl, spin_limit);
    return hipGetLastError();
}

int batch_lines_per_tile(int logm)
{
    switch (logm) {
#define MDFT_BLPW_CASE(L) \
    case L: return LineCfg<L, BATCH_THREADS>::LPW;
        MDFT_BLPW_CASE(1)
        MDFT_BLPW_CASE(2)
        MDFT_BLPW_CASE(3)
        MDFT_BLPW_CASE(4)
        MDFT_BLPW_CASE(5)
        MDFT_BLPW_CASE(6)
        MDFT_BLPW_CASE(7)
        MDFT_BLPW_CASE(8)
        MDFT_BLPW_CASE(9)
        MDFT_BLPW_CASE(10)
        MDFT_BLPW_CASE(11)
        MDFT_BLPW_CASE(12)
        MDFT_BLPW_CASE(13)
#undef MDFT_BLPW_CASE
    default: return 0;
    }
}

hipError_t batch_occupancy(int* blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch, BATCH_THREADS, 0);
}

int batch_threads() { return BATCH_THREADS; }
int batch_max_logm() { return 13; }

}  // namespace mdft
