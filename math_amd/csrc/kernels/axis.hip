// Axis-pass kernels: one launch transforms every line of one axis of one
// image (the reference records rows_to_buffer, FFT, mul_d, FFT,
// rows_from_buffer as five dispatches per axis, src/gpu/dft/dft.cpp:326-347;
// here they are one fused kernel per axis).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"
#include "launch.hpp"

namespace mdft {

// Workgroup sizes of the single-image kernels.  Narrow: 256 threads, or
// one line when a line needs more (M = 8192 or 6144 -> 512, M = 16384 -> 1024).
// Wide: 1024 threads, so a tile holds 16384 points (e.g. 4 columns of 4096),
// which widens the contiguous run a column tile reads from each row.
__host__ __device__ constexpr int axis_threads(int shape, bool wide)
{
    const int tpl = m_of(shape) / ept_of(shape);
    const int base = wide ? 1024 : 256;
    return tpl > base ? tpl : base;
}

template <int S, int T, bool BLUE, bool COLS>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(TILE_WAVES))) void k_axis(AxisJob job)
{
    constexpr int LF = LineCfg<S, T>::template lds_floats<COLS>();
    __shared__ float lds[LF > 0 ? LF : 1];
    const int tile = COLS ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    line_tile<S, T, BLUE, COLS>(job, tile, lds, threadIdx.x);
}

template <int S, int T>
static hipError_t launch_axis_tt(const AxisJob& job, bool blue, bool cols, hipStream_t stream)
{
    constexpr int LPW = LineCfg<S, T>::LPW;
    // in 64 bits: count may be up to 2^31 - 1
    const int64_t grid64 = (int64_t(job.count) + LPW - 1) / LPW;
    if (job.count < 0 || grid64 > 0x7fffffff) return hipErrorInvalidValue;
    if (grid64 == 0) return hipSuccess;
    const unsigned grid = (unsigned)grid64;
    if (blue) {
        if constexpr (S >= 2) {
            if (cols)
                hipLaunchKernelGGL((k_axis<S, T, true, true>), dim3(grid), dim3(T), 0, stream, job);
            else
                hipLaunchKernelGGL((k_axis<S, T, true, false>), dim3(grid), dim3(T), 0, stream, job);
        } else {
            return hipErrorInvalidValue;
        }
    } else {
        if (cols)
            hipLaunchKernelGGL((k_axis<S, T, false, true>), dim3(grid), dim3(T), 0, stream, job);
        else
            hipLaunchKernelGGL((k_axis<S, T, false, false>), dim3(grid), dim3(T), 0, stream, job);
    }
    return hipGetLastError();
}

// threads: 1024 = wide tiles, 256 = narrow, 128 / 64 = small tiles for
// passes with few lines (clamped up to one line per workgroup)
template <int S>
static hipError_t launch_axis_t(const AxisJob& job, bool blue, bool cols, int threads, hipStream_t stream)
{
    constexpr int TPL = m_of(S) / ept_of(S);
    constexpr int TN = axis_threads(S, false), TW = axis_threads(S, true);
    if (threads >= 1024 && TW != TN) return launch_axis_tt<S, TW>(job, blue, cols, stream);
    if constexpr (TPL <= 64) {
        if (threads <= 64) return launch_axis_tt<S, 64>(job, blue, cols, stream);
    }
    if constexpr (TPL <= 128) {
        if (threads <= 128) return launch_axis_tt<S, 128>(job, blue, cols, stream);
    }
    return launch_axis_tt<S, TN>(job, blue, cols, stream);
}

// one word written, so the context loads this file's kernels at creation
__global__ void k_warm_axis(uint32_t* word)
{
    if (threadIdx.x == 0) word[0] = 0u;
}

hipError_t warm_axis(uint32_t* word, hipStream_t stream)
{
    hipLaunchKernelGGL(k_warm_axis, dim3(1), dim3(64), 0, stream, word);
    return hipGetLastError();
}

hipError_t launch_axis(int shape, const AxisJob& job, bool blue, bool cols, int threads, hipStream_t stream)
{
    switch (shape) {
    case 1: return launch_axis_t<1>(job, blue, cols, threads, stream);
    case 2: return launch_axis_t<2>(job, blue, cols, threads, stream);
    case 3: return launch_axis_t<3>(job, blue, cols, threads, stream);
    case 4: return launch_axis_t<4>(job, blue, cols, threads, stream);
    case 5: return launch_axis_t<5>(job, blue, cols, threads, stream);
    case 6: return launch_axis_t<6>(job, blue, cols, threads, stream);
    case 7: return launch_axis_t<7>(job, blue, cols, threads, stream);
    case 8: return launch_axis_t<8>(job, blue, cols, threads, stream);
    case 9: return launch_axis_t<9>(job, blue, cols, threads, stream);
    case 10: return launch_axis_t<10>(job, blue, cols, threads, stream);
    case 11: return launch_axis_t<11>(job, blue, cols, threads, stream);
    case 12: return launch_axis_t<12>(job, blue, cols, threads, stream);
    case 13: return launch_axis_t<13>(job, blue, cols, threads, stream);
    case 14: return launch_axis_t<14>(job, blue, cols, threads, stream);
    case F3_BASE + 6: return launch_axis_t<F3_BASE + 6>(job, blue, cols, threads, stream);
    case F3_BASE + 7: return launch_axis_t<F3_BASE + 7>(job, blue, cols, threads, stream);
    case F3_BASE + 8: return launch_axis_t<F3_BASE + 8>(job, blue, cols, threads, stream);
    case F3_BASE + 9: return launch_axis_t<F3_BASE + 9>(job, blue, cols, threads, stream);
    case F3_BASE + 10: return launch_axis_t<F3_BASE + 10>(job, blue, cols, threads, stream);
    case F3_BASE + 11: return launch_axis_t<F3_BASE + 11>(job, blue, cols, threads, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mdft
