// Batched 2-D DFTs of mixed sizes.
//
// A batch is a list of work items over a table of axis jobs (one row job
// and one column job per image).  An item is one tile of one job, or a whole
// small image: a workgroup runs the row tile, makes its stores visible to its
// own waves, then runs the column tile, with no cross-workgroup
// synchronisation at all.  Three launch shapes share the item code:
//
//  * k_batch_grid: one workgroup per item, items laid out by the host so
//    each job runs on one XCD; a batch is two launches, whole images + row
//    tiles (256-thread tiles), then column tiles (512-thread tiles, 2-8
//    columns each for padded lengths 1024-4096, so each row of a tile is a
//    longer contiguous run, with two workgroups per CU).
//  * k_batch_persistent: a grid sized to the device's residency takes items
//    from a ticket counter, one launch per phase.
//  * k_batch_persistent with dependencies: one launch for everything; a
//    column tile waits on a per-image counter of finished row tiles (a
//    device-scope release/acquire pair per tile, which costs an L2
//    write-back per row tile).
//
// Every line runs the same per-line arithmetic as the single-image kernels
// (same radix schedule, explicit fmas), so a result is bitwise the same alone
// and in any mix.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"
#include "launch.hpp"

namespace mdft {

// longest padded length a tile of `threads` threads holds (one line per tile at most 8192 points)
__host__ __device__ constexpr int batch_maxlogm(int threads)
{
    return (threads >= 1024 ? 10 : threads >= 512 ? 9 : 8) + fft_bits(13) > 13
               ? 13
               : (threads >= 1024 ? 10 : threads >= 512 ? 9 : 8) + fft_bits(13);
}

template <int S, int T>
constexpr int lds_need()
{
    if constexpr (m_of(S) / ept_of(S) > T) {
        return 0;
    } else {
        constexpr int a = LineCfg<S, T>::template lds_floats<false>();
        constexpr int b = LineCfg<S, T>::template lds_floats<true>();
        return a > b ? a : b;
    }
}

constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int T>
constexpr int batch_lds()
{
    return cmax(cmax(cmax(cmax(lds_need<5, T>(), lds_need<6, T>()), cmax(lds_need<7, T>(), lds_need<8, T>())),
                     cmax(cmax(cmax(lds_need<9, T>(), lds_need<10, T>()), cmax(lds_need<11, T>(), lds_need<12, T>())),
                          lds_need<13, T>())),
                cmax(cmax(cmax(lds_need<F3_BASE + 6, T>(), lds_need<F3_BASE + 7, T>()),
                          cmax(lds_need<F3_BASE + 8, T>(), lds_need<F3_BASE + 9, T>())),
                     cmax(lds_need<F3_BASE + 10, T>(), lds_need<F3_BASE + 11, T>())));
}

template <int S, int T>
__device__ __forceinline__ void run_tile_body(const AxisJob& job, int blue, int cols, int tile, float* lds)
{
    if constexpr (m_of(S) / ept_of(S) <= T) {
        if constexpr (S >= 2) {
            if (blue) {
                if (cols)
                    line_tile<S, T, true, true>(job, tile, lds, threadIdx.x);
                else
                    line_tile<S, T, true, false>(job, tile, lds, threadIdx.x);
                return;
            }
        }
        if (cols)
            line_tile<S, T, false, true>(job, tile, lds, threadIdx.x);
        else
            line_tile<S, T, false, false>(job, tile, lds, threadIdx.x);
    }
}

// Grid kernels inline every variant (register allocation per region,
// occupancy-constrained); the persistent kernels call them out of line, since
// inlining into the ticket loop keeps too much alive across it.
template <int S, int T, bool INL>
struct RunTile;
template <int S, int T>
struct RunTile<S, T, true> {
    __device__ __forceinline__ static void run(const AxisJob& job, int blue, int cols, int tile, float* lds)
    {
        run_tile_body<S, T>(job, blue, cols, tile, lds);
    }
};
template <int S, int T>
__device__ __noinline__ void run_tile_call(const AxisJob& job, int blue, int cols, int tile, float* lds)
{
    run_tile_body<S, T>(job, blue, cols, tile, lds);
}
template <int S, int T>
struct RunTile<S, T, false> {
    __device__ __forceinline__ static void run(const AxisJob& job, int blue, int cols, int tile, float* lds)
    {
        run_tile_call<S, T>(job, blue, cols, tile, lds);
    }
};

template <int T, bool INL>
__device__ __forceinline__ void run_job_tile(const BatchJob& bj, int tile, float* lds)
{
    switch (bj.shape) {
    case 1: RunTile<1, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 2: RunTile<2, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 3: RunTile<3, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 4: RunTile<4, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 5: RunTile<5, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 6: RunTile<6, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 7: RunTile<7, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 8: RunTile<8, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 9: RunTile<9, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 10: RunTile<10, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 11: RunTile<11, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 12: RunTile<12, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case 13: RunTile<13, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case F3_BASE + 6: RunTile<F3_BASE + 6, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case F3_BASE + 7: RunTile<F3_BASE + 7, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case F3_BASE + 8: RunTile<F3_BASE + 8, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case F3_BASE + 9: RunTile<F3_BASE + 9, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case F3_BASE + 10: RunTile<F3_BASE + 10, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    case F3_BASE + 11: RunTile<F3_BASE + 11, T, INL>::run(bj.job, bj.blue, bj.cols, tile, lds); break;
    default: break;
    }
}

// One item; every branch is on a wave-uniform value.
template <int T, bool INL>
__device__ __forceinline__ void run_item(const BatchJob* __restrict__ jobs, uint2 item, float* lds)
{
    const uint32_t x = __builtin_amdgcn_readfirstlane(item.x);
    const uint32_t y = __builtin_amdgcn_readfirstlane(item.y);
    const BatchJob bj = jobs[x];
    if (y & BATCH_WHOLE) {
        run_job_tile<T, INL>(bj, 0, lds);
        // the column tile reads what this workgroup's waves just stored
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const BatchJob bc = jobs[y & ~BATCH_WHOLE];
        run_job_tile<T, INL>(bc, 0, lds);
    } else {
        run_job_tile<T, INL>(bj, (int)y, lds);
    }
}

// Four waves per SIMD (<= 128 VGPRs): 256-thread tiles (4 workgroups per
// CU), 512-thread tiles (two) or 1024-thread tiles (one).
template <int T>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(TILE_WAVES))) void k_batch_grid(
    const BatchJob* __restrict__ jobs, const uint2* __restrict__ items)
{
    __shared__ float lds[batch_lds<T>()];
    // the host lays items out per XCD (block b runs on XCD b % 8)
    run_item<T, true>(jobs, items[blockIdx.x], lds);
}

// The item loop contains barriers, so its control flow is kept free of
// lane-divergent regions: every branch tests a value broadcast with
// readfirstlane or a wave index, and single-lane atomics are issued by the
// whole wave with an increment of 1 on lane 0 and 0 elsewhere.  (A
// thread-0-only region inside the loop lets the compiler rotate the loop so
// the other lanes reach the next barrier without passing the ticket update.)
// Deadlock freedom of the dependency mode: an item waits only on items with
// smaller tickets, taken by running workgroups that never wait on later ones.
template <int T>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(TILE_WAVES))) void k_batch_persistent(
    const BatchJob* __restrict__ jobs, const uint2* __restrict__ items, uint32_t nitems, uint32_t* ctl,
    uint32_t spin_limit)
{
    __shared__ float lds[batch_lds<T>()];
    __shared__ uint32_t s_item;
    uint32_t* ticket = ctl;
    uint32_t* err = ctl + 1;
    uint32_t* counters = ctl + BATCH_CTL_HEADER;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x) >> 6;
    const uint32_t one_on_lane0 = (threadIdx.x & 63) == 0 ? 1u : 0u;
    for (;;) {
        if (wave == 0) {
            const uint32_t t = atomicAdd(ticket, one_on_lane0);
            s_item = __builtin_amdgcn_readfirstlane(t);
        }
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= nitems) break;
        const uint2 item = items[it];
        const BatchJob bj = jobs[__builtin_amdgcn_readfirstlane(item.x)];
        if (bj.dep >= 0) {
            // every wave polls the image's row counter itself, then acquires
            uint32_t spins = 0;
            for (;;) {
                const uint32_t c = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(&counters[bj.dep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (c >= (uint32_t)bj.dep_need) break;
                if (++spins > spin_limit) {
                    atomicOr(err, one_on_lane0);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        run_item<T, false>(jobs, item, lds);
        if (bj.sig >= 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // release this wave's stores
            __syncthreads();
            if (wave == 0) atomicAdd(&counters[bj.sig], one_on_lane0);
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t value)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_batch_grid(int threads, const BatchJob* jobs, const uint2* items, uint32_t nitems,
                             hipStream_t stream)
{
    if (nitems == 0) return hipSuccess;
    if (threads == 256)
        hipLaunchKernelGGL(k_batch_grid<256>, dim3(nitems), dim3(256), 0, stream, jobs, items);
    else if (threads == 512)
        hipLaunchKernelGGL(k_batch_grid<512>, dim3(nitems), dim3(512), 0, stream, jobs, items);
    else
        hipLaunchKernelGGL(k_batch_grid<1024>, dim3(nitems), dim3(1024), 0, stream, jobs, items);
    return hipGetLastError();
}

hipError_t warm_batch(uint32_t* word, hipStream_t stream)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, word, 1u, 0u);
    return hipGetLastError();
}

hipError_t launch_batch_persistent(int threads, const BatchJob* jobs, const uint2* items, uint32_t nitems,
                                   uint32_t* ctl, uint32_t nctl, int grid, uint32_t spin_limit, hipStream_t stream)
{
    if (nitems == 0) return hipSuccess;
    // clear ticket, error flag and counters with a kernel (graph-capture safe)
    hipLaunchKernelGGL(k_fill_u32, dim3((nctl + 255) / 256), dim3(256), 0, stream, ctl, nctl, 0u);
    if (threads == 256)
        hipLaunchKernelGGL(k_batch_persistent<256>, dim3(grid), dim3(256), 0, stream, jobs, items, nitems, ctl,
                           spin_limit);
    else if (threads == 512)
        hipLaunchKernelGGL(k_batch_persistent<512>, dim3(grid), dim3(512), 0, stream, jobs, items, nitems, ctl,
                           spin_limit);
    else
        hipLaunchKernelGGL(k_batch_persistent<1024>, dim3(grid), dim3(1024), 0, stream, jobs, items, nitems, ctl,
                           spin_limit);
    return hipGetLastError();
}

int batch_lines_per_tile(int threads, int shape)
{
    if (shape < 1 || (!is_f3(shape) && shape > batch_maxlogm(threads)) ||
        (is_f3(shape) && (f3_k(shape) < F3_MIN_K || f3_k(shape) > F3_MAX_K)))
        return 0;
    const int tpl = m_of(shape) / ept_of(shape);
    return tpl <= threads ? threads / tpl : 0;
}

hipError_t batch_occupancy(int threads, int* blocks_per_cu)
{
    if (threads == 256)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch_persistent<256>, 256, 0);
    if (threads == 512)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch_persistent<512>, 512, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_batch_persistent<1024>, 1024, 0);
}

int batch_max_logm(int threads) { return batch_maxlogm(threads); }

}  // namespace mdft
