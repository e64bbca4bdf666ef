// Image path of the view (src/gpu/dft/compute.cpp:91-200): luminance of an
// RGBA rectangle -> forward 2-D DFT -> |X|/(w*h) into an R32F image.  The
// reference runs copy_input, the ten DFT dispatches and copy_output; here the
// luminance read is fused into the first axis pass and the magnitude write
// into the last, so the complex image is written and read once in between.
// Lines longer than one tile use the unfused kernels below around the
// ordinary transform.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"
#include "launch.hpp"

namespace mdft {

__host__ __device__ constexpr int image_threads(int shape, bool wide)
{
    const int tpl = m_of(shape) / ept_of(shape);
    const int base = wide ? 1024 : 256;
    return tpl > base ? tpl : base;
}

template <int S, int T, bool BLUE, bool COLS, int IO>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(TILE_WAVES))) void k_image_axis(AxisJob job, ImageIO io)
{
    constexpr int LF = LineCfg<S, T>::template lds_floats<COLS>();
    __shared__ float lds[LF > 0 ? LF : 1];
    const int tile = COLS ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    line_tile<S, T, BLUE, COLS, IO>(job, tile, lds, threadIdx.x, io);
}

template <int S, int T, bool COLS, int IO>
static hipError_t launch_image_tt(const AxisJob& job, const ImageIO& io, bool blue, hipStream_t stream)
{
    constexpr int LPW = LineCfg<S, T>::LPW;
    const int64_t grid64 = (int64_t(job.count) + LPW - 1) / LPW;
    if (job.count < 0 || grid64 > 0x7fffffff) return hipErrorInvalidValue;
    if (grid64 == 0) return hipSuccess;
    const unsigned grid = (unsigned)grid64;
    if (blue) {
        if constexpr (S >= 2)
            hipLaunchKernelGGL((k_image_axis<S, T, true, COLS, IO>), dim3(grid), dim3(T), 0, stream, job, io);
        else
            return hipErrorInvalidValue;
    } else {
        hipLaunchKernelGGL((k_image_axis<S, T, false, COLS, IO>), dim3(grid), dim3(T), 0, stream, job, io);
    }
    return hipGetLastError();
}

// rows: IO_LUM, or IO_LUM|IO_MAG for a one-row image; columns: IO_MAG, or
// IO_LUM|IO_MAG for a one-column image.  Wide tiles only for the column
// pass that stores magnitudes (the shape of the ordinary column pass).
template <int S>
static hipError_t launch_image_t(const AxisJob& job, const ImageIO& io, bool blue, bool cols, int mode, bool wide,
                                 hipStream_t stream)
{
    constexpr int TN = image_threads(S, false), TW = image_threads(S, true);
    if (!cols) {
        if (mode == IO_LUM) return launch_image_tt<S, TN, false, IO_LUM>(job, io, blue, stream);
        if (mode == (IO_LUM | IO_MAG)) return launch_image_tt<S, TN, false, IO_LUM | IO_MAG>(job, io, blue, stream);
        return hipErrorInvalidValue;
    }
    if (mode == IO_MAG) {
        if (wide && TW != TN) return launch_image_tt<S, TW, true, IO_MAG>(job, io, blue, stream);
        return launch_image_tt<S, TN, true, IO_MAG>(job, io, blue, stream);
    }
    if (mode == (IO_LUM | IO_MAG)) return launch_image_tt<S, TN, true, IO_LUM | IO_MAG>(job, io, blue, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_image_axis(int shape, const AxisJob& job, const ImageIO& io, bool blue, bool cols, int mode,
                             bool wide, hipStream_t stream)
{
    switch (shape) {
    case 1: return launch_image_t<1>(job, io, blue, cols, mode, wide, stream);
    case 2: return launch_image_t<2>(job, io, blue, cols, mode, wide, stream);
    case 3: return launch_image_t<3>(job, io, blue, cols, mode, wide, stream);
    case 4: return launch_image_t<4>(job, io, blue, cols, mode, wide, stream);
    case 5: return launch_image_t<5>(job, io, blue, cols, mode, wide, stream);
    case 6: return launch_image_t<6>(job, io, blue, cols, mode, wide, stream);
    case 7: return launch_image_t<7>(job, io, blue, cols, mode, wide, stream);
    case 8: return launch_image_t<8>(job, io, blue, cols, mode, wide, stream);
    case 9: return launch_image_t<9>(job, io, blue, cols, mode, wide, stream);
    case 10: return launch_image_t<10>(job, io, blue, cols, mode, wide, stream);
    case 11: return launch_image_t<11>(job, io, blue, cols, mode, wide, stream);
    case 12: return launch_image_t<12>(job, io, blue, cols, mode, wide, stream);
    case 13: return launch_image_t<13>(job, io, blue, cols, mode, wide, stream);
    case 14: return launch_image_t<14>(job, io, blue, cols, mode, wide, stream);
    case F3_BASE + 6: return launch_image_t<F3_BASE + 6>(job, io, blue, cols, mode, wide, stream);
    case F3_BASE + 7: return launch_image_t<F3_BASE + 7>(job, io, blue, cols, mode, wide, stream);
    case F3_BASE + 8: return launch_image_t<F3_BASE + 8>(job, io, blue, cols, mode, wide, stream);
    case F3_BASE + 9: return launch_image_t<F3_BASE + 9>(job, io, blue, cols, mode, wide, stream);
    case F3_BASE + 10: return launch_image_t<F3_BASE + 10>(job, io, blue, cols, mode, wide, stream);
    case F3_BASE + 11: return launch_image_t<F3_BASE + 11>(job, io, blue, cols, mode, wide, stream);
    default: return hipErrorInvalidValue;
    }
}

// ---- unfused kernels (long lines, and the 1 x 1 image) --------------------

// luminance of the w x h rectangle into a complex image (imaginary part 0);
// a grid-stride loop over the w*h pixels
__global__ void k_image_luminance(ImageIO io, float2* dst, int w, int64_t npix)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t y = i / w, x = i - y * w;
        dst[i] = make_float2(image_luminance(io, y * io.ils + x * io.iel), 0.0f);
    }
}

// |X| * scale of a complex image into the R32F output
__global__ void k_image_magnitude(ImageIO io, const float2* src, int w, int64_t npix)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t y = i / w, x = i - y * w;
        io.mag[y * io.mls + x * io.mel] = cabs_scaled(src[i], io.mscale);
    }
}

static unsigned pixel_grid(int64_t npix)
{
    const int64_t g = (npix + 255) / 256;
    return (unsigned)(g < 65536 ? g : 65536);
}

// one word written, so the context loads this file's kernels at creation
__global__ void k_warm_image(uint32_t* word)
{
    if (threadIdx.x == 0) word[0] = 0u;
}

hipError_t warm_image(uint32_t* word, hipStream_t stream)
{
    hipLaunchKernelGGL(k_warm_image, dim3(1), dim3(64), 0, stream, word);
    return hipGetLastError();
}

hipError_t launch_image_luminance(const ImageIO& io, float2* dst, int w, int h, hipStream_t stream)
{
    const int64_t npix = int64_t(w) * h;
    if (npix <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_image_luminance, dim3(pixel_grid(npix)), dim3(256), 0, stream, io, dst, w, npix);
    return hipGetLastError();
}

hipError_t launch_image_magnitude(const ImageIO& io, const float2* src, int w, int h, hipStream_t stream)
{
    const int64_t npix = int64_t(w) * h;
    if (npix <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_image_magnitude, dim3(pixel_grid(npix)), dim3(256), 0, stream, io, src, w, npix);
    return hipGetLastError();
}

}  // namespace mdft
