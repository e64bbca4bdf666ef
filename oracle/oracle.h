/*
 * CPU oracle for the 2-D DFT path.  TEST INFRASTRUCTURE ONLY: this library is
 * the checker.  Only tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it; the product library never links or calls it.
 *
 * Three parts:
 *   1. The reference's host-side plan math restated (compute_m, compute_h,
 *      compute_h2, shared_size), following src/gpu/dft/function.cpp:32-118.
 *   2. An FP32 restatement of the reference's GPU pipeline
 *      (src/gpu/dft/dft.cpp:55-136,262-347, fft.cpp:207-330 and the GLSL
 *      kernels dft_mul.comp, dft_mul_d.comp, dft_fft_shared.comp,
 *      dft_fft_global.comp, dft_bit_reverse.comp), threaded over lines.  It is
 *      also the CPU baseline when FFTW is absent.
 *   3. An FP64 oracle of the 2-D DFT itself (direct DFT for short lines,
 *      radix-2 FFT for powers of two, Bluestein otherwise) with the reference's
 *      semantics: forward unnormalised, inverse scaled by 1/(w*h)
 *      (src/dft/fftw.cpp:129-159).
 *
 * Complex arrays are interleaved (re, im), row-major, element (x, y) at
 * index y*w + x, width the fast axis (src/gpu/dft/code/dft_mul.comp:94).
 */
#ifndef MDFT_ORACLE_H
#define MDFT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* function.cpp:32-45.  Returns -1 where the reference raises
 * "Error size <n> for compute m" (n < 1 or n > 2^29). */
int orc_compute_m(int n);

/* function.cpp:49-71.  h[l] = coef*exp(+-i*pi*l^2/n), l < n, the exponent
 * reduced in 64-bit integers.  h: 2*n doubles. */
void orc_compute_h(int n, int inverse, double coef, double* h);

/* function.cpp:75-92 (with the errata of compute.cpp:26-43).  h2: 2*m doubles. */
void orc_compute_h2(int n, int m, const double* h, double* h2);

/* function.cpp:94-105 for complex<float> elements. */
unsigned orc_shared_size(unsigned dft_size, unsigned max_shared_memory_bytes);

/* FP32 restatement of the reference's batched 1-D radix-2 FFT
 * (fft.cpp:207-330): count lines of length n (a power of two) in place.
 * The inverse is normalised by 1/n the way the reference does it (division by
 * N_MAX in the LDS kernel, x0.5 per global stage). */
void orc_fft32(int count, int n, float* data, int inverse, unsigned shared_size);

/* FP32 restatement of the reference's chirp multiply, dft_mul.comp:44-97. */
void orc_mul_data32(float* x, long long l, int n, int inverse);

/* The four FP32 diagonals of dft.cpp:77-111: d1 tables have m1 entries,
 * d2 tables m2 entries (2*m floats each). */
void orc_diagonals32(int n1, int n2, unsigned lds_bytes, float* d1_fwd, float* d1_inv, float* d2_fwd,
                     float* d2_inv);

/* FP32 restatement of the whole reference 2-D pipeline, dft.cpp:326-347, in
 * place on data (2*w*h floats).  lds_bytes is the Vulkan
 * maxComputeSharedMemorySize the reference plans its FFT tile with
 * (65536 gives the 8192-point tile).  nthreads <= 0 uses one thread. */
int orc_dft2d_ref32(int w, int h, int inverse, float* data, unsigned lds_bytes, int nthreads);

/* FP64 1-D DFT of one line, reference semantics (inverse scaled by 1/n).
 * in and out: 2*n doubles, may not alias. */
void orc_dft1d_64(int n, int inverse, const double* in, double* out);

/* FP64 direct O(n^2) DFT of one line with exact integer reduction of j*k mod n. */
void orc_dft1d_direct64(int n, int inverse, const double* in, double* out);

/* FP64 2-D DFT, reference semantics.  in/out: 2*w*h doubles (may alias). */
int orc_dft2d_64(int w, int h, int inverse, const double* in, double* out, int nthreads);

/* Same, from FP32 input, upcast exactly. */
int orc_dft2d_64_from32(int w, int h, int inverse, const float* in, double* out, int nthreads);

/* Error of an FP32 result against an FP64 reference over count complex
 * values.  out[0] = ||got-ref||_2/||ref||_2, out[1] = max|got-ref| /
 * max|ref|, out[2] = max|got-ref|, out[3] = sum|got-ref|/sum|got| (the
 * reference's relative L1 discrepancy, src/gpu/dft/test/test_dft.cpp:63-93). */
void orc_error_metrics(size_t count, const float* got, const double* ref, double* out);

/* Counter-based synthetic data: dist 0 = uniform[-1,1), dist 1 = N(0,1).
 * Deterministic for (seed, index); 2*count floats. */
void orc_fill(uint64_t seed, int dist, size_t count, float* out);

/* Image path of the view (src/gpu/dft/compute.cpp:91-200): the luminance
 * 0.2126 r + 0.7152 g + 0.0722 b of each pixel of the w x h rectangle at
 * (x, y) of an RGBA image with a row pitch of `pitch` pixels
 * (dft_copy_input.comp:37-54; fmt 0 = 4 floats, fmt 1 = 4 x uint8 UNORM),
 * rounded in the order fma(0.0722, b, fma(0.7152, g, 0.2126 * r)), into
 * lum (w*h floats). */
void orc_image_luminance(int fmt, const void* src, int64_t pitch, int x, int y, int w, int h, float* lum);

#ifdef __cplusplus
}
#endif

#endif
