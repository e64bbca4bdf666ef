// Lines longer than one tile (padded length M > 16384): four-step FFT through
// a scratch buffer, M = M1 * M2 with M1, M2 <= 16384 (the reference runs a
// bit reversal, an LDS tile and log2(M/tile) global radix-2 stages here,
// src/gpu/dft/fft.cpp:242-262, dft_fft_global.comp:44-78).
//
// With n = n1 + M1*n2 and k = k2 + M2*k1:
//   step A: for every (line, n1): length-M2 FFT over n2, times w_M^(n1*k2),
//           stored contiguously per (line, n1) in scratch;
//   step C: for every (line, k2): length-M1 FFT over n1, stored at k2 + M2*k1
//           (natural order).
// A power-of-two line is A then C.  A Bluestein line is three steps: A (chirp
// and zero pad on the load); B, in place on the scratch, per k2: the length-M1
// FFT over n1 (C's), times D and conjugated, then the first half of the
// second FFT with the roles of M1 and M2 exchanged (k = k2 + M2*k1 and
// j = j1 + M1*j2: w_M^(k*j) = w_M2^(k2*j2) * w_M^(k2*j1) * w_M1^(k1*j1)), i.e. a
// length-M1 FFT over k1 times w_M^(k2*j1); and C', per j1: the length-M2 FFT
// over k2 into j1 + M1*j2 (conjugate, chirp, crop, 1/n on the store).  Every
// step is a strided tile of whole sub-lines run with the column
// (lines-fastest) mapping.
//
// Padded lengths 2^29 and 2^30 (M2 * M3 above one tile) take three factors,
// n = n1 + M1*(n2 + M2*n3), k = kr + Mr*k1 with Mr = M2*M3, kr = k3 + M3*k2:
//   A1: per (line, n1, n2) the M3-point FFT over n3, times w_Mr^(n2*k3);
//   A2: in place, per (line, n1, k3) the M2-point FFT over n2 (the Mr-point
//       FFT of every n1, its output kr spread as k3 + M1*M3*k2 beside n1);
//   C:  per (line, kr) times w_M^(n1*kr) on the load, the M1-point FFT over n1,
//       stored at kr + Mr*k1.
// A Bluestein line runs A1 (chirp) and A2, then C in place fused like B: the
// M1-point FFT, conj(v * D), the M1-point FFT over k1 and w_M^(kr*j1), with j1
// in the slot of n1; the second transform's Mr-point FFT over kr then runs as
//   X1: in place, per (line, j1, k3) the M2-point FFT over k2, times w_Mr^(k3*jr2);
//   X2: per (line, j1, jr2) the M3-point FFT over k3 (contiguous), stored at
//       j = j1 + M1*(jr2 + M2*jr3) with the chirp and crop.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"
#include "launch.hpp"

namespace mdft {

__device__ __forceinline__ int64_t low_bits(int64_t v, int lg) { return v & ((int64_t(1) << lg) - 1); }

// line L of a step maps to (L / sub) * ls + r * ls2 with r = L % sub, element
// p to p * es; r may split once more (sub2 > 0): (r % sub2) * ls2 + (r / sub2) * ls3.
// sub and sub2 are powers of two (lg: their log2, lg2 = -1 without the split).
__device__ __forceinline__ int64_t line_base3(int64_t L, int lg, int64_t ls, int64_t ls2, int lg2, int64_t ls3)
{
    const int64_t r = low_bits(L, lg);
    const int64_t off = lg2 >= 0 ? low_bits(r, lg2) * ls2 + (r >> lg2) * ls3 : r * ls2;
    return (L >> lg) * ls + off;
}

// w_M^e, e < M, from the two-level table
__device__ __forceinline__ float2 large_twiddle(const LargeJob& job, int64_t e)
{
    return cmul(job.tw_lo[e & (LARGE_LO - 1)], job.tw_hi[e >> LARGE_LO_BITS]);
}

template <int LOGM, int T, bool FUSED>
__global__ __launch_bounds__(T) void k_large_step(LargeJob job)
{
    using C = LineCfg<LOGM, T>;
    constexpr int LF = C::template lds_floats<true>();
    __shared__ float lds[LF > 0 ? LF : 1];
    const int tid = threadIdx.x;
    const int ll = tid % C::LPW;
    const int t = tid / C::LPW;
    const int64_t line = int64_t(xcd_remap(blockIdx.x, gridDim.x)) * C::LPW + ll;
    const bool valid = line < job.count;
    const uint32_t sgn = job.inverse ? 0x80000000u : 0u;
    // sub-line coordinate inside the long line (the chirp load's element offset)
    const int64_t sub_idx = low_bits(line, job.lg_sub_in);
    // twiddle exponent per element of this sub-line
    const int64_t tq = (low_bits(line, job.lg_tmod) >> job.lg_tdiv) * job.tmul;

    float2 v[C::EPT];
    {
        const float2* s =
            job.src + (valid ? line_base3(line, job.lg_sub_in, job.ls_in, job.ls2_in, job.lg_sub2_in, job.ls3_in) : 0);
#pragma unroll
        for (int i = 0; i < C::EPT; ++i) {
            const int64_t pos = t + i * C::TPL;
            float2 x = make_float2(0.0f, 0.0f);
            if (valid) {
                if (job.load_mode == LARGE_LOAD_CHIRP) {
                    // element n1 + M1*n2 of the Bluestein input: chirped, zero past n
                    const int64_t j = sub_idx + job.m1 * pos;
                    if (j < job.n) x = cmul(s[pos * job.es_in], cxor(job.chirp[j], sgn));
                } else if (job.load_mode == LARGE_LOAD_CONJ) {
                    x = cxor(s[pos * job.es_in], sgn);
                } else if (job.load_mode == LARGE_LOAD_TWIDDLE) {
                    x = cmul(s[pos * job.es_in], large_twiddle(job, tq * pos));
                } else {
                    x = s[pos * job.es_in];
                }
            }
            v[i] = x;
        }
    }

    FftTile<LOGM, T, true>::run(v, t, ll, lds, job.tw);
    const int64_t q_line = low_bits(line, job.lg_sub_out);
    if constexpr (FUSED) {
        // conj(v * D) at k = k2 + M2*k1 (q_line = k2 < M2, pos = k1 < M1; three factors: kr + Mr*k1;
        // in bounds on every line), then the next length-M1 FFT over k1; its outputs j1 take the twiddle below
#pragma unroll
        for (int i = 0; i < C::EPT; ++i) {
            const int64_t k = q_line + job.kmul * (t + i * C::TPL);
            v[i] = cmul_conj(v[i], cxor(job.dtab[k], sgn));
        }
        FftTile<LOGM, T, true>::run(v, t, ll, lds, job.tw);
    }

    if (!valid) return;
    float2* o = job.dst + line_base3(line, job.lg_sub_out, job.ls_out, job.ls2_out, job.lg_sub2_out, job.ls3_out);
#pragma unroll
    for (int i = 0; i < C::EPT; ++i) {
        const int64_t pos = t + i * C::TPL;
        float2 y = v[i];
        if (job.store_mode == LARGE_STORE_TWIDDLE || job.store_mode == LARGE_STORE_DIAG_FFT) {
            // the step's twiddle w_M^(tq*pos), tq*pos < M: n1*k2 (A), k2*j1 (B), M1*n2*k3 (A1),
            // kr*j1 (fused C), M1*k3*jr2 (X1)
            y = cmul(y, large_twiddle(job, tq * pos));
        } else if (job.store_mode == LARGE_STORE_DIAG) {
            // natural index k = k2 + M2*k1; forward conj(v*D), inverse conj(v*conj(D))
            const int64_t k = q_line + job.kmul * pos;
            y = cmul_conj(y, cxor(job.dtab[k], sgn));
        } else if (job.store_mode == LARGE_STORE_CHIRP) {
            const int64_t k = q_line + job.kmul * pos;
            if (k >= job.n) continue;
            y = cscale(cmul_conj(y, cxor(job.chirp[k], sgn ^ 0x80000000u)), job.scale);
        } else if (job.store_mode == LARGE_STORE_CONJ) {
            y = cscale(cxor(y, sgn), job.scale);
        }
        o[pos * job.es_out] = y;
    }
}

// log2 of a power of two, -1 for anything else
static int lg_of(int64_t v)
{
    if (v < 1 || (v & (v - 1)) != 0) return -1;
    int r = 0;
    while ((int64_t(1) << r) < v) ++r;
    return r;
}

template <int LOGM, int T>
static hipError_t launch_large_tt(const LargeJob& job0, hipStream_t stream)
{
    constexpr int LPW = LineCfg<LOGM, T>::LPW;
    const int64_t grid = (job0.count + LPW - 1) / LPW;
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    // the kernel splits sub-line indices with shifts and masks
    LargeJob job = job0;
    job.lg_sub_in = lg_of(job.sub_in);
    job.lg_sub_out = lg_of(job.sub_out);
    job.lg_sub2_in = job.sub2_in > 0 ? lg_of(job.sub2_in) : -1;
    job.lg_sub2_out = job.sub2_out > 0 ? lg_of(job.sub2_out) : -1;
    if (job.lg_sub_in < 0 || job.lg_sub_out < 0 || job.sub2_in < 0 || job.sub2_out < 0 ||
        (job.sub2_in > 0 && job.lg_sub2_in < 0) || (job.sub2_out > 0 && job.lg_sub2_out < 0))
        return hipErrorInvalidValue;
    const bool twiddled = job.load_mode == LARGE_LOAD_TWIDDLE || job.store_mode == LARGE_STORE_TWIDDLE ||
                          job.store_mode == LARGE_STORE_DIAG_FFT;
    if (twiddled) {
        job.lg_tmod = lg_of(job.tmod);
        job.lg_tdiv = lg_of(job.tdiv);
        if (job.lg_tmod < 0 || job.lg_tdiv < 0 || job.tmul < 1) return hipErrorInvalidValue;
    } else {
        job.lg_tmod = 0, job.lg_tdiv = 0;  // exponent 0
    }
    if (job.store_mode == LARGE_STORE_DIAG_FFT)
        hipLaunchKernelGGL((k_large_step<LOGM, T, true>), dim3((unsigned)grid), dim3(T), 0, stream, job);
    else
        hipLaunchKernelGGL((k_large_step<LOGM, T, false>), dim3((unsigned)grid), dim3(T), 0, stream, job);
    return hipGetLastError();
}

// Tiles of 1024 threads (most sub-lines per tile, the widest contiguous runs
// per strided row) unless that leaves fewer than min_blocks workgroups; then
// the largest of 256 and 64 threads that does not, else the smallest tile
// that holds a sub-line (a few long lines: every CU gets work).
template <int LOGM>
static hipError_t launch_large_t(const LargeJob& job, int min_blocks, hipStream_t stream)
{
    constexpr int TPL = m_of(LOGM) / ept_of(LOGM);
    auto blocks = [&](int t) { return (job.count + t / TPL - 1) / (t / TPL); };
    if constexpr (TPL <= 256) {
        if (blocks(1024) < min_blocks) {
            if constexpr (TPL <= 64) {
                if (blocks(256) < min_blocks) return launch_large_tt<LOGM, 64>(job, stream);
            }
            return launch_large_tt<LOGM, 256>(job, stream);
        }
    }
    return launch_large_tt<LOGM, 1024>(job, stream);
}

// one word written, so the context loads this file's kernels at creation
__global__ void k_warm_large(uint32_t* word)
{
    if (threadIdx.x == 0) word[0] = 0u;
}

hipError_t warm_large(uint32_t* word, hipStream_t stream)
{
    hipLaunchKernelGGL(k_warm_large, dim3(1), dim3(64), 0, stream, word);
    return hipGetLastError();
}

hipError_t launch_large_step(int logm, const LargeJob& job, int min_blocks, hipStream_t stream)
{
    switch (logm) {
    case 1: return launch_large_t<1>(job, min_blocks, stream);
    case 2: return launch_large_t<2>(job, min_blocks, stream);
    case 3: return launch_large_t<3>(job, min_blocks, stream);
    case 4: return launch_large_t<4>(job, min_blocks, stream);
    case 5: return launch_large_t<5>(job, min_blocks, stream);
    case 6: return launch_large_t<6>(job, min_blocks, stream);
    case 7: return launch_large_t<7>(job, min_blocks, stream);
    case 8: return launch_large_t<8>(job, min_blocks, stream);
    case 9: return launch_large_t<9>(job, min_blocks, stream);
    case 10: return launch_large_t<10>(job, min_blocks, stream);
    case 11: return launch_large_t<11>(job, min_blocks, stream);
    case 12: return launch_large_t<12>(job, min_blocks, stream);
    case 13: return launch_large_t<13>(job, min_blocks, stream);
    case 14: return launch_large_t<14>(job, min_blocks, stream);
    default: return hipErrorInvalidValue;
    }
}

// exp(-2*pi*i*j/M) for j < 2^LARGE_LO_BITS, and exp(-2*pi*i*j*2^LARGE_LO_BITS/M)
// for j < M / 2^LARGE_LO_BITS, in FP64 then rounded
__global__ void k_gen_large_twiddles(float2* lo, float2* hi, int64_t m)
{
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nhi = m >> LARGE_LO_BITS;
    if (j < LARGE_LO) {
        double s, c;
        sincospi(2.0 * (double)j / (double)m, &s, &c);
        lo[j] = make_float2((float)c, (float)-s);
    }
    if (j < nhi) {
        double s, c;
        sincospi(2.0 * (double)(j << LARGE_LO_BITS) / (double)m, &s, &c);
        hi[j] = make_float2((float)c, (float)-s);
    }
}

hipError_t gen_large_twiddles(float2* lo, float2* hi, int64_t m, hipStream_t stream)
{
    const int64_t nhi = m >> LARGE_LO_BITS;
    const int64_t n = nhi > LARGE_LO ? nhi : LARGE_LO;
    hipLaunchKernelGGL(k_gen_large_twiddles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, lo, hi, m);
    return hipGetLastError();
}

}  // namespace mdft
