// Plan-time table generation on the device, in FP64 with exactly reduced
// integer exponents, rounded once to FP32:
//   * chirp[k]   = exp(-i*pi*(k^2 mod 2n)/n), k < n   (src/gpu/dft/function.cpp:49-71)
//   * dtab[q]    = (1/M) * sum_l h2[l] exp(-2*pi*i*(l*q mod M)/M), the FFT of the
//                  circulant embedding h2 of conj(chirp)  (function.cpp:75-92,
//                  dft.cpp:77-111), with the 1/M of the inverse FFT folded in
//   * twiddles   = per-pass Stockham rows: w^1..w^4 of w = exp(-2*pi*i*k/L), exponents k*e mod L
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"

namespace mdft {

// exp(-2*pi*i*q/den) for integers 0 <= q < den, den <= 2^62
__device__ __forceinline__ double2 unit_root(int64_t q, int64_t den)
{
    double s, c;
    sincospi(2.0 * (double)q / (double)den, &s, &c);
    return make_double2(c, -s);
}

// row k of a pass: w^1..w^TW_ROW of w = exp(-2*pi*i*k/(ns*radix))
__global__ void k_gen_twiddle_block(float2* out, int ns, int radix)
{
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t len = int64_t(ns) * radix;
    if (idx >= int64_t(ns) * TW_ROW) return;
    const int64_t k = idx / TW_ROW, e = idx % TW_ROW + 1;
    const double2 w = unit_root((k * e) % len, len);
    out[idx] = make_float2((float)w.x, (float)w.y);
}

// chirp c[k] in FP64 and FP32; with dir != nullptr also the direction tables
// of the tile kernels (layout of dir_tables_len): load c / conj(c), store
// conj(c) / c/n (the inverse's 1/n folded in, rounded once from FP64)
__global__ void k_gen_chirp(double2* chirp64, float2* chirp32, float2* dir, int n, int m)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t two_n = 2 * int64_t(n);
    // k^2 mod 2n without overflow for k < 2^31: reduce k first
    const int64_t kr = k % two_n;
    const int64_t q = (int64_t)(((unsigned __int128)kr * kr) % (unsigned __int128)two_n);
    const double2 c = unit_root(q, two_n);
    chirp64[k] = c;
    const float2 c32 = make_float2((float)c.x, (float)c.y);
    chirp32[k] = c32;
    if (dir) {
        const int64_t out = 2 * int64_t(n) + 2 * int64_t(m);
        dir[k] = c32;
        dir[n + k] = make_float2(c32.x, -c32.y);
        dir[out + k] = make_float2(c32.x, -c32.y);
        dir[out + n + k] = make_float2((float)(c.x / (double)n), (float)(c.y / (double)n));
    }
}

// the diagonal and its conjugate into the direction tables
__global__ void k_dir_diagonal(float2* dir, const float2* dtab, int n, int m)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const float2 d = dtab[q];
    dir[2 * int64_t(n) + q] = d;
    dir[2 * int64_t(n) + m + q] = make_float2(d.x, -d.y);
    dir[4 * int64_t(n) + 2 * int64_t(m) + q] = make_float2(0.0f, 0.0f);  // tail read past the store blocks
}

// cos(2*pi*q/m), q < m
__global__ void k_gen_cos(double* cosm, int m)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= m) return;
    cosm[q] = cospi(2.0 * (double)q / (double)m);
}

// The circulant kernel is even (h2[l] == h2[m-l]), so its DFT is
//   D[q] = h2[0] + sum_{0<l<n} f_l * h2[l] * cos(2*pi*l*q/m),
// f_l = 2 except f = 1 for l == m/2 (its own mirror when m == 2n-2).
// h2[l] = conj(chirp[l]).  One workgroup per output, lanes split the sum.
__global__ void __launch_bounds__(256) k_gen_dtab(float2* dtab, const double2* chirp64, const double* cosm, int n,
                                                  int m)
{
    const int q = blockIdx.x;
    double sr = 0.0, si = 0.0;
    for (int l = 1 + threadIdx.x; l < n; l += blockDim.x) {
        const double f = (2 * l == m) ? 1.0 : 2.0;
        const int64_t idx = (int64_t(l) * q) % m;
        const double c = cosm[idx] * f;
        sr += chirp64[l].x * c;
        si -= chirp64[l].y * c;
    }
    __shared__ double red[2][256];
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double inv_m = 1.0 / (double)m;
        const double re = (red[0][0] + chirp64[0].x) * inv_m;
        const double im = (red[1][0] - chirp64[0].y) * inv_m;
        dtab[q] = make_float2((float)re, (float)im);
    }
}

// Circulant embedding of conj(chirp) rounded to FP32, for lengths whose
// diagonal is computed by an FFT instead of the direct sum.
__global__ void k_gen_h2_f32(float2* h2, const double2* chirp64, int64_t n, int64_t m)
{
    const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= m) return;
    double2 c = make_double2(0.0, 0.0);
    if (l < n)
        c = chirp64[l];
    else if (l > m - n)
        c = chirp64[m - l];
    h2[l] = make_float2((float)c.x, (float)-c.y);
}

// --- host launchers --------------------------------------------------------

hipError_t gen_h2_f32(float2* h2, const double2* chirp64, int64_t n, int64_t m, hipStream_t stream)
{
    hipLaunchKernelGGL(k_gen_h2_f32, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, h2, chirp64, n, m);
    return hipGetLastError();
}

hipError_t gen_twiddles(float2* table, hipStream_t stream)
{
    auto shape_rows = [&](int shape) {
        for (int p = 1; p < npass_of(shape); ++p) {
            const int ns = ns_of(shape, p);
            const int radix = radix_of(shape, p);
            const int64_t len = int64_t(ns) * TW_ROW;
            float2* out = table + tw_off(shape) + tw_pass_off(shape, p);
            const int blocks = (int)((len + 255) / 256);
            hipLaunchKernelGGL(k_gen_twiddle_block, dim3(blocks), dim3(256), 0, stream, out, ns, radix);
        }
    };
    for (int logm = 1; logm <= MAX_LOGM; ++logm) shape_rows(logm);
    for (int k = F3_MIN_K; k <= F3_MAX_K; ++k) shape_rows(F3_BASE + k);
    return hipGetLastError();
}

hipError_t gen_chirp(double2* chirp64, float2* chirp32, float2* dir, int n, int m, hipStream_t stream)
{
    hipLaunchKernelGGL(k_gen_chirp, dim3((unsigned)((int64_t(n) + 255) / 256)), dim3(256), 0, stream, chirp64, chirp32,
                       dir, n, m);
    return hipGetLastError();
}

hipError_t gen_dir_diagonal(float2* dir, const float2* dtab, int n, int m, hipStream_t stream)
{
    hipLaunchKernelGGL(k_dir_diagonal, dim3((unsigned)((int64_t(m) + 255) / 256)), dim3(256), 0, stream, dir, dtab, n,
                       m);
    return hipGetLastError();
}

hipError_t gen_dtab(float2* dtab, const double2* chirp64, double* cos_scratch, int n, int m, hipStream_t stream)
{
    hipLaunchKernelGGL(k_gen_cos, dim3((m + 255) / 256), dim3(256), 0, stream, cos_scratch, m);
    hipLaunchKernelGGL(k_gen_dtab, dim3(m), dim3(256), 0, stream, dtab, chirp64, (const double*)cos_scratch, n, m);
    return hipGetLastError();
}

}  // namespace mdft
