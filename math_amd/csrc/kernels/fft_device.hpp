// Device-side building blocks of the 2-D DFT line passes (gfx950 / CDNA4).
//
// One "axis pass" transforms `count` lines of logical length n, laid out with
// element stride es and line stride ls inside a row-major complex<float>
// image (rows: es = 1, ls = w; columns: es = w, ls = 1).  A line of length n
// is transformed through an FFT of length M = 2^LOGM: directly when n is a
// power of two (M = n), otherwise by Bluestein's chirp-z method with M the
// reference's padded length (src/gpu/dft/function.cpp:32-45) and the chirp
// tables of function.cpp:49-92 generated in FP64 on the device.
//
// A workgroup owns LPW whole lines (a "tile").  Each thread holds EPT = 16
// points of one line in VGPRs, at positions t + i*TPL (TPL = M/16 threads per
// line).  The FFT is a radix-16 Stockham autosort (passes of 16 and one
// remainder pass), DIT form: pass p gathers its operands from LDS at the
// thread's fixed positions, multiplies by the per-pass twiddle row, runs a
// radix-R butterfly in registers and scatters to the Stockham positions.  The
// first pass needs no twiddles and reads straight from HBM; the last pass
// leaves its results at the thread's own positions, so it writes straight to
// HBM and Bluestein's two FFTs chain without an LDS exchange.  LDS holds one
// exchange buffer, split into real and imaginary planes with one pad word per
// 16 to keep the scatter conflict-free.
//
// Only the forward FFT exists: an inverse transform is conj(FFT(conj(x))),
// folded into the load, the D multiply and the store.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdft {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Products are written with explicit fused multiply-adds and the library is
// built with -ffp-contract=off, so every rounding is fixed by the source and a
// line gives bitwise the same result in every kernel that runs it.
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, a.y * b.y), __builtin_fmaf(a.y, b.x, -(a.x * b.y)));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// conj(a * b)
__device__ __forceinline__ float2 cmul_conj(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(-a.x, b.y, -(a.y * b.x)));
}
// flip the sign of the imaginary part when sgn is 0x80000000 (conjugate), keep it when 0
__device__ __forceinline__ float2 cxor(float2 a, uint32_t sgn)
{
    return make_float2(a.x, __uint_as_float(__float_as_uint(a.y) ^ sgn));
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// ---------------------------------------------------------------------------
// compile-time line geometry
// ---------------------------------------------------------------------------
// Bits per radix pass of a line of padded length 2^logm: radix-16 passes (16
// points per thread) or radix-8 passes (8 points per thread, half the
// registers and LDS per thread).  A function of the length alone, so a line
// runs the same schedule in every kernel.  16384-point lines keep radix 16 so
// that one line fits a 1024-thread workgroup.
#ifndef MDFT_FFT_BITS
#define MDFT_FFT_BITS 4
#endif
static_assert(MDFT_FFT_BITS == 3 || MDFT_FFT_BITS == 4, "radix-8 or radix-16 passes");
__host__ __device__ constexpr int fft_bits(int logm) { return logm >= 14 ? 4 : MDFT_FFT_BITS; }
// minimum waves per SIMD the tile kernels are compiled for (register cap 512 / waves)
#ifndef MDFT_TILE_WAVES
#define MDFT_TILE_WAVES 4
#endif
__host__ __device__ constexpr int ept_of(int logm)
{
    return logm < fft_bits(logm) ? (1 << logm) : (1 << fft_bits(logm));
}
__host__ __device__ constexpr int npass_of(int logm) { return (logm + fft_bits(logm) - 1) / fft_bits(logm); }
__host__ __device__ constexpr int pass_bits(int logm, int p)
{
    return (logm - fft_bits(logm) * p) >= fft_bits(logm) ? fft_bits(logm) : (logm - fft_bits(logm) * p);
}

// Twiddle rows of pass p >= 1 (Ns = (2^fft_bits)^p before the pass, radix R): row k
// (k < Ns) holds TW_ROW = 4 entries w^1, w^2, w^3, w^4 of
// w = exp(-2*pi*i*k/(Ns*R)) (32 bytes, four rows per cache line); the other
// powers are products of these (expand_twiddles).
constexpr int TW_ROW = 4;
__host__ __device__ constexpr int64_t tw_pass_len(int logm, int p)
{
    return (pass_bits(logm, p) > 0 ? (int64_t(1) << (fft_bits(logm) * p)) : 0) * TW_ROW;
}
__host__ __device__ constexpr int64_t tw_pass_off(int logm, int p)
{
    int64_t off = 0;
    for (int q = 1; q < p; ++q) off += tw_pass_len(logm, q);
    return off;
}
__host__ __device__ constexpr int64_t tw_len(int logm)
{
    int64_t s = 0;
    for (int p = 1; p < npass_of(logm); ++p) s += tw_pass_len(logm, p);
    return s;
}
// All LOGM in [1, MAX_LOGM] share one table; offset of LOGM's block.
constexpr int MAX_LOGM = 14;
__host__ __device__ constexpr int64_t tw_off(int logm)
{
    int64_t s = 0;
    for (int l = 1; l < logm; ++l) s += tw_len(l);
    return s;
}
__host__ __device__ constexpr int64_t tw_total() { return tw_off(MAX_LOGM + 1); }

// LDS word of line position p in the exchange written by a pass of stride NS.
// A pass writes its outputs NS positions apart and the next pass reads
// positions TPL apart.  Radix-16 passes: when a 32-lane half-wave covers 32
// consecutive positions of one line (ONE_LINE), the reads are conflict-free
// unpadded and the writes need: NS = 1 (lanes 16 apart) one pad word per 32,
// NS = 16 (lanes in two runs of 16, 256 apart) 16 pad words per 256, NS >= 256
// (consecutive lanes) none; otherwise lanes span several lines and one pad
// word per 16 with an odd line stride keeps them apart.  Radix-8 passes: one
// pad word per 8.  Every map satisfies f(t + i*TPL) = f(t) + f(i*TPL) for
// t < TPL, so LDS offsets are immediates.
template <int NS, bool ONE_LINE, int B>
__host__ __device__ constexpr int xpad(int p)
{
    if constexpr (B == 3) {
        return p + (p >> 3);
    } else if constexpr (!ONE_LINE) {
        return p + (p >> 4);
    } else if constexpr (NS == 1) {
        return p + (p >> 5);
    } else if constexpr (NS == 16) {
        return p + 16 * (p >> 8);
    } else {
        return p;
    }
}

template <int LOGM, int THREADS>
struct LineCfg {
    static constexpr int M = 1 << LOGM;
    static constexpr int EPT = ept_of(LOGM);
    static constexpr int TPL = M / EPT;          // threads per line
    static constexpr int LPW = THREADS / TPL;    // lines per tile
    static constexpr int NPASS = npass_of(LOGM);
    static_assert(TPL <= THREADS, "line does not fit the workgroup");
    static_assert(LPW * TPL == THREADS, "workgroup must hold whole lines");
    static constexpr int B = fft_bits(LOGM);
    static constexpr int PADM = M + (M >> B);
    // a 32-lane half-wave covers 32 consecutive positions of one line
    template <bool COLS>
    static constexpr bool one_line()
    {
        return COLS ? LPW == 1 : TPL >= 32;
    }
    // Line stride in LDS words.  Lanes of one wave cover q consecutive
    // positions of 32/q lines (rows: q = TPL; columns, lines fastest:
    // q = 32/LPW); a stride of q times an odd number puts them on distinct
    // banks.
    template <bool COLS>
    static constexpr int lstride()
    {
        const int q = COLS ? (LPW >= 32 ? 1 : 32 / LPW) : (TPL >= 32 ? 0 : TPL);
        if (q == 0) return PADM;
        int s = PADM;
        while (s % (2 * q) != q) ++s;
        return s;
    }
    template <bool COLS>
    static constexpr int lds_floats()
    {
        return NPASS > 1 ? 2 * LPW * lstride<COLS>() : 0;
    }
};

// ---------------------------------------------------------------------------
// radix-R forward DFT in registers (R <= 16), natural order in and out
// ---------------------------------------------------------------------------
// cos(2*pi*m/16), sin(2*pi*m/16)
__device__ constexpr float kC16[16] = {1.0f,
                                       0.92387953251128675613f,
                                       0.70710678118654752440f,
                                       0.38268343236508977173f,
                                       0.0f,
                                       -0.38268343236508977173f,
                                       -0.70710678118654752440f,
                                       -0.92387953251128675613f,
                                       -1.0f,
                                       -0.92387953251128675613f,
                                       -0.70710678118654752440f,
                                       -0.38268343236508977173f,
                                       0.0f,
                                       0.38268343236508977173f,
                                       0.70710678118654752440f,
                                       0.92387953251128675613f};
__device__ constexpr float kS16[16] = {0.0f,
                                       0.38268343236508977173f,
                                       0.70710678118654752440f,
                                       0.92387953251128675613f,
                                       1.0f,
                                       0.92387953251128675613f,
                                       0.70710678118654752440f,
                                       0.38268343236508977173f,
                                       0.0f,
                                       -0.38268343236508977173f,
                                       -0.70710678118654752440f,
                                       -0.92387953251128675613f,
                                       -1.0f,
                                       -0.92387953251128675613f,
                                       -0.70710678118654752440f,
                                       -0.38268343236508977173f};

// d * exp(-2*pi*i*m/16) for a compile-time m
template <int m>
__device__ __forceinline__ float2 twc(float2 d)
{
    if constexpr (m == 0) {
        return d;
    } else if constexpr (m == 4) {
        return make_float2(d.y, -d.x);
    } else if constexpr (m == 2) {
        const float r = 0.70710678118654752440f;
        return make_float2(r * (d.x + d.y), r * (d.y - d.x));
    } else if constexpr (m == 6) {
        const float r = 0.70710678118654752440f;
        return make_float2(r * (d.y - d.x), -r * (d.x + d.y));
    } else {
        const float c = kC16[m], s = kS16[m];
        return make_float2(__builtin_fmaf(d.x, c, d.y * s), __builtin_fmaf(d.y, c, -(d.x * s)));
    }
}

__host__ __device__ constexpr int bitrev_small(int v, int bits)
{
    int r = 0;
    for (int i = 0; i < bits; ++i) {
        r = (r << 1) | (v & 1);
        v >>= 1;
    }
    return r;
}

// one radix-2 DIF stage of span HALF over R points.  ZHI (first stage
// only): u[R/2 + 1 ..] are known to be zero, so every pair but the first
// is (a, 0) and becomes (a, twiddle * a) without additions.
template <int R, int HALF, int BASE, int K, bool ZHI>
__device__ __forceinline__ void dif_bfly(float2* u)
{
    if constexpr (BASE < R) {
        if constexpr (K < HALF) {
            const float2 a = u[BASE + K];
            if constexpr (ZHI && K > 0) {
                u[BASE + K + HALF] = twc<K*(16 / (2 * HALF))>(a);
            } else {
                const float2 b = u[BASE + K + HALF];
                u[BASE + K] = cadd(a, b);
                u[BASE + K + HALF] = twc<K*(16 / (2 * HALF))>(csub(a, b));
            }
            dif_bfly<R, HALF, BASE, K + 1, ZHI>(u);
        } else {
            dif_bfly<R, HALF, BASE + 2 * HALF, 0, ZHI>(u);
        }
    }
}

template <int R, int HALF, bool ZHI>
__device__ __forceinline__ void dif_stages(float2* u)
{
    if constexpr (HALF >= 1) {
        dif_bfly<R, HALF, 0, 0, ZHI>(u);
        dif_stages<R, HALF / 2, false>(u);
    }
}

// ZHI: inputs u[R/2 + 1 ..] are zero (the zero-padded half of a Bluestein
// line in its first pass)
template <int R, bool ZHI = false>
__device__ __forceinline__ void dft_fwd(float2* u)
{
    static_assert(R >= 1 && R <= 16 && (R & (R - 1)) == 0, "radix");
    if constexpr (R > 1) {
        dif_stages<R, R / 2, ZHI>(u);
        constexpr int bits = (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : 4;
        float2 t[R];
#pragma unroll
        for (int i = 0; i < R; ++i) t[i] = u[bitrev_small(i, bits)];
#pragma unroll
        for (int i = 0; i < R; ++i) u[i] = t[i];
    }
}

// ---------------------------------------------------------------------------
// the job of one axis pass
// ---------------------------------------------------------------------------
struct AxisJob {
    const float2* src;     // input image base
    float2* dst;           // output image base (may equal src)
    int64_t es;            // element stride (complex elements)
    int64_t ls;            // line stride
    int n;                 // logical line length
    int count;             // number of lines
    int inverse;           // 0 forward, 1 inverse
    float scale;           // 1/n for the inverse, 1 otherwise
    // Bluestein tables of the job's direction (c = exp(-i*pi*k^2/n)):
    const float2* chirp;      // load: c (forward) or conj(c) (inverse), k < n
    const float2* dtab;       // middle: D or conj(D), D = FFT_M(conj(c) embedded) / M
    const float2* chirp_out;  // store: conj(c) (forward) or c/n (inverse), k < n
    const float2* tw;      // twiddle block of this LOGM
};

// ---------------------------------------------------------------------------
// image input and output of the view path: the luminance of an RGBA pixel
// becomes the complex input (src/gpu/dft/code/dft_copy_input.comp:37-54) and
// |X| * scale the R32F output (dft_copy_output.comp:34-48)
// ---------------------------------------------------------------------------
enum : int { IO_LUM = 1, IO_MAG = 2 };
enum : int { PIXEL_RGBA32F = 0, PIXEL_RGBA8 = 1 };
struct ImageIO {
    const void* img;  // RGBA pixel at the rectangle's origin
    int64_t iel;      // pixel stride along a line
    int64_t ils;      // pixel stride between lines
    int fmt;          // PIXEL_*
    float mscale;     // 1/(w*h)
    float* mag;       // R32F output at the rectangle's origin
    int64_t mel;      // float stride along a line
    int64_t mls;      // float stride between lines
};

// 0.2126 r + 0.7152 g + 0.0722 b (dft_copy_input.comp:37-40), fixed rounding order
__host__ __device__ __forceinline__ float luminance_rgb(float r, float g, float b)
{
    return __builtin_fmaf(0.0722f, b, __builtin_fmaf(0.7152f, g, 0.2126f * r));
}

__device__ __forceinline__ float image_luminance(const ImageIO& io, int64_t idx)
{
    if (io.fmt == PIXEL_RGBA8) {
        const uchar4 c = static_cast<const uchar4*>(io.img)[idx];
        // UNORM: c / 255
        return luminance_rgb((float)c.x / 255.0f, (float)c.y / 255.0f, (float)c.z / 255.0f);
    }
    const float4 c = static_cast<const float4*>(io.img)[idx];
    return luminance_rgb(c.x, c.y, c.z);
}

__device__ __forceinline__ float cabs_scaled(float2 y, float s)
{
    return __builtin_sqrtf(__builtin_fmaf(y.x, y.x, y.y * y.y)) * s;
}

// ---------------------------------------------------------------------------
// FFT of the tile held in registers
// ---------------------------------------------------------------------------
// Barrier around an LDS exchange.  When every line of a tile lies inside one
// wave (row mapping with TPL <= 64), an exchange never crosses waves, and the
// LDS operations of one wave are performed in order, so a wave-level
// code-motion barrier replaces the workgroup barrier: waves of the tile then
// run their passes independently.
template <bool WAVE>
__device__ __forceinline__ void exchange_sync()
{
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Twiddle rows of pass P for the thread's butterflies are loaded one pass
// ahead (pass 1's with the tile's input), so the table latency overlaps the
// LDS exchange instead of following it.  Off for 16384 points, where the
// extra registers spill.
template <int LOGM>
__host__ __device__ constexpr bool tw_prefetch()
{
    return LOGM <= 13;
}

// float4s of a row a radix-R butterfly reads: (w^1, w^2) and (w^3, w^4)
template <int R>
__host__ __device__ constexpr int tw_quads()
{
    return R <= 2 ? 1 : 2;
}

template <int LOGM, int THREADS, int P>
struct TwLoad {
    using C = LineCfg<LOGM, THREADS>;
    static constexpr int R = 1 << pass_bits(LOGM, P);
    static constexpr int Q = tw_quads<R>();
    static constexpr int NB = C::EPT / R;
    static constexpr int NS = 1 << (C::B * P);
    static_assert(NB * Q <= C::EPT / 2, "twiddle registers");
    __device__ __forceinline__ static void run(float4 (&w)[C::EPT / 2], int t, const float2* tw)
    {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int k = (t + b * C::TPL) & (NS - 1);
            const float4* row = reinterpret_cast<const float4*>(tw + tw_pass_off(LOGM, P) + int64_t(k) * TW_ROW);
#pragma unroll
            for (int r2 = 0; r2 < Q; ++r2) w[b * Q + r2] = row[r2];
        }
    }
};

// the R twiddles of one butterfly (wt[0] unused) from its row: w^1..w^4
// loaded, w^(4a+b) = w^(4a) * w^b, w^8 = w^4 * w^4, w^12 = w^8 * w^4 (at most
// four roundings from the FP64-generated entries)
template <int R>
__device__ __forceinline__ void expand_twiddles(float2 (&wt)[R], const float4* q)
{
    wt[1] = make_float2(q[0].x, q[0].y);
    if constexpr (R > 2) {
        wt[2] = make_float2(q[0].z, q[0].w);
        wt[3] = make_float2(q[1].x, q[1].y);
    }
    if constexpr (R > 4) {
        wt[4] = make_float2(q[1].z, q[1].w);
#pragma unroll
        for (int b = 1; b < 4; ++b) wt[4 + b] = cmul(wt[4], wt[b]);
    }
    if constexpr (R > 8) {
        wt[8] = cmul(wt[4], wt[4]);
        wt[12] = cmul(wt[8], wt[4]);
#pragma unroll
        for (int b = 1; b < 4; ++b) {
            wt[8 + b] = cmul(wt[8], wt[b]);
            wt[12 + b] = cmul(wt[12], wt[b]);
        }
    }
}

// ZIN (first pass only): slots EPT/2 + 1 .. EPT of every thread are zero.
// PF: the first PF slots of a table (pfs, indexed by line position) are
// loaded into pf one pass before the last, so their latency overlaps the
// last exchange and butterflies; slots PF .. PFL-1 at the start of the last
// pass, after its exchange (with one pass, all PFL at its start).
template <int LOGM, int THREADS, bool COLS, int P, bool ZIN = false, int PF = 0, int PFL = 0>
struct FftPass {
    using C = LineCfg<LOGM, THREADS>;
    static constexpr bool WAVE = !COLS && C::TPL <= 64;
    static constexpr int RB = pass_bits(LOGM, P);
    static constexpr int R = 1 << RB;
    static constexpr int NB = C::EPT / R;
    static constexpr int NS = 1 << (C::B * P);
    static constexpr int RADIX = 1 << C::B;
    static constexpr bool LAST = (P == C::NPASS - 1);
    static constexpr int LST = C::template lstride<COLS>();
    static constexpr bool ONE = C::template one_line<COLS>();

    __device__ __forceinline__ static void run(float2 (&v)[C::EPT], int t, float* re, float* im, const float2* tw,
                                               float4 (&w)[C::EPT / 2], const float2* pfs, float2 (&pf)[C::EPT])
    {
        static_assert(PF <= C::EPT && PFL <= C::EPT, "prefetch slots");
        if constexpr (PF > 0 && C::NPASS == 1) {
#pragma unroll
            for (int i = 0; i < (PFL > PF ? PFL : PF); ++i) pf[i] = pfs[t + i * C::TPL];
        }
        if constexpr (P > 0) {
            exchange_sync<WAVE>();
            // layout of the previous pass (stride NS/16); the offset of slot i
            // is a compile-time constant (an LDS immediate)
            const int q0 = xpad<NS / RADIX, ONE, C::B>(t);
#pragma unroll
            for (int i = 0; i < C::EPT; ++i) {
                const int q = q0 + xpad<NS / RADIX, ONE, C::B>(i * C::TPL);
                v[i] = make_float2(re[q], im[q]);
            }
            if constexpr (LAST && PFL > PF) {
#pragma unroll
                for (int i = PF; i < PFL; ++i) pf[i] = pfs[t + i * C::TPL];
            }
        }
        // butterfly b reads and writes only the slots b + NB*r: update in place
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * C::TPL;
            const int k = j & (NS - 1);
            float2 u[R];
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = v[b + NB * r];
            if constexpr (P > 0) {
                constexpr int Q = tw_quads<R>();
                float4 q[2];
                const float4* row = reinterpret_cast<const float4*>(tw + tw_pass_off(LOGM, P) + int64_t(k) * TW_ROW);
#pragma unroll
                for (int r2 = 0; r2 < Q; ++r2) q[r2] = tw_prefetch<LOGM>() ? w[b * Q + r2] : row[r2];
                float2 wt[R];
                expand_twiddles<R>(wt, q);
#pragma unroll
                for (int r = 1; r < R; ++r) u[r] = cmul(u[r], wt[r]);
            }
            static_assert(!ZIN || (P == 0 && NB == 1), "zero inputs only in a single-butterfly first pass");
            dft_fwd<R, ZIN>(u);
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + NB * r] = u[r];
        }
        if constexpr (!LAST) {
            if constexpr (tw_prefetch<LOGM>()) TwLoad<LOGM, THREADS, P + 1>::run(w, t, tw);
            if constexpr (PF > 0 && P + 2 == C::NPASS) {
#pragma unroll
                for (int i = 0; i < PF; ++i) pf[i] = pfs[t + i * C::TPL];
            }
            exchange_sync<WAVE>();
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int j = t + b * C::TPL;
                const int k = j & (NS - 1);
                const int base = (j - k) * R + k;
                // NS == 1 (first pass, full radix): base is a multiple of the
                // radix, so the outputs are contiguous after padding; NS >= the
                // radix: the outputs lie in one padding block apart from whole
                // blocks, so each step adds xpad(NS)
                static_assert(NS == 1 ? R == RADIX : NS % RADIX == 0, "linear write offsets");
                constexpr int step = xpad<NS, ONE, C::B>(NS);
                const int q0 = xpad<NS, ONE, C::B>(base);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int q = q0 + r * step;
                    re[q] = v[b + NB * r].x;
                    im[q] = v[b + NB * r].y;
                }
            }
            FftPass<LOGM, THREADS, COLS, P + 1, false, PF, PFL>::run(v, t, re, im, tw, w, pfs, pf);
        }
    }
};

// ZIN: the upper half of the line is zero apart from position M/2, so slots
// EPT/2 + 1 .. EPT of every thread hold zeros (the caller set them to
// constant zero).  PF > 0: pf[i] = pfs[t + i*TPL] for i < max(PF, PFL) on
// return, loaded during the FFT (pfs must be readable at every such
// position).
template <int LOGM, int THREADS, bool COLS, bool ZIN = false, int PF = 0, int PFL = 0>
struct FftTile {
    using C = LineCfg<LOGM, THREADS>;
    __device__ __forceinline__ static void run(float2 (&v)[C::EPT], int t, int ll, float* lds, const float2* tw)
    {
        static_assert(PF == 0 && PFL == 0, "prefetch needs a destination");
        float2 pf[C::EPT];
        run(v, t, ll, lds, tw, nullptr, pf);
    }
    __device__ __forceinline__ static void run(float2 (&v)[C::EPT], int t, int ll, float* lds, const float2* tw,
                                               const float2* pfs, float2 (&pf)[C::EPT])
    {
        constexpr int LST = C::template lstride<COLS>();
        float* re = lds + ll * LST;
        float* im = lds + C::LPW * LST + ll * LST;
        float4 w[C::EPT / 2];
        if constexpr (tw_prefetch<LOGM>() && C::NPASS > 1) TwLoad<LOGM, THREADS, 1>::run(w, t, tw);
        FftPass<LOGM, THREADS, COLS, 0, ZIN, PF, PFL>::run(v, t, re, im, tw, w, pfs, pf);
    }
};

// ---------------------------------------------------------------------------
// one tile of one axis pass: load (+chirp), FFT, [x D, FFT], store (+chirp)
// ---------------------------------------------------------------------------
// IO (image kernels only): IO_LUM loads luminance from io.img instead of
// complex values from job.src; IO_MAG stores |X| * io.mscale to io.mag
// instead of complex values to job.dst.
template <int LOGM, int THREADS, bool BLUE, bool COLS, int IO = 0>
__device__ __forceinline__ void line_tile(const AxisJob& job, int tile, float* lds, int tid, ImageIO io = ImageIO{})
{
    using C = LineCfg<LOGM, THREADS>;
    const int ll = COLS ? (tid % C::LPW) : (tid / C::TPL);
    const int t = COLS ? (tid / C::LPW) : (tid % C::TPL);
    const int line = tile * C::LPW + ll;
    const bool valid = line < job.count;
    const int n = job.n;
    // rows are contiguous (stride 1, known at compile time); a column steps by
    // the image width.  Consecutive slots of a thread are TPL elements apart.
    const int64_t es = COLS ? job.es : 1;
    const int64_t step = int64_t(C::TPL) * es;
    // direction enters as a sign mask: conjugation is one xor, and the
    // inverse's 1/n is a multiply by job.scale (exactly 1 for the forward)
    const uint32_t sgn = job.inverse ? 0x80000000u : 0u;
    // a power-of-two line fills the FFT (n == M); only Bluestein pads
    auto in_line = [&](int pos) { return !BLUE || pos < n; };
    // Bluestein: n <= M/2 + 1, so slot i holds a position of the line only
    // for i <= EPT/2 (slot EPT/2 only at t = 0 when n = M/2 + 1); the other
    // slots are zero on input and not stored on output.
    constexpr int LIVE = BLUE ? C::EPT / 2 + 1 : C::EPT;

    float2 v[C::EPT];
    {
        const float2* s = job.src + int64_t(valid ? line : 0) * job.ls + int64_t(t) * es;
#pragma unroll
        for (int i = 0; i < C::EPT; ++i) {
            const uint32_t pos = uint32_t(t + i * C::TPL);
            float2 x = make_float2(0.0f, 0.0f);
            if (i < LIVE && valid && in_line((int)pos)) {
                if constexpr ((IO & IO_LUM) != 0)
                    x = make_float2(image_luminance(io, int64_t(line) * io.ils + int64_t(pos) * io.iel), 0.0f);
                else
                    x = s[i * step];
                // forward x*c, inverse x*conj(c) (direction tables); power of two: conj for the inverse
                x = BLUE ? cmul(x, job.chirp[pos]) : cxor(x, sgn);
            }
            v[i] = x;
        }
    }

    // Bluestein: D arrives during the first FFT, its first half one pass
    // before the end and the rest at the start of the last pass (up to 4096
    // points; earlier loads of all of D, or D of 8192 points, spill at the
    // 128-VGPR cap), and the live slots of the output chirp during the second
    constexpr int PFD = (BLUE && C::NPASS <= 3) ? C::EPT / 2 : 0;
    constexpr int PFDL = PFD > 0 ? C::EPT : 0;
    float2 pf[C::EPT];
    FftTile<LOGM, THREADS, COLS, BLUE, PFD, PFDL>::run(v, t, ll, lds, job.tw, job.dtab, pf);

    if constexpr (BLUE) {
        // forward conj(v*D), inverse conj(v)*D = conj(v*conj(D)); the outer
        // conjugation turns the next forward FFT into the inverse one
#pragma unroll
        for (int i = 0; i < C::EPT; ++i) {
            const uint32_t pos = uint32_t(t + i * C::TPL);
            v[i] = cmul_conj(v[i], i < PFDL ? pf[i] : job.dtab[pos]);
        }
        // the output chirp block is followed by readable entries up to
        // position M/2 + M/16 (dir_tables_len), so every live slot can be loaded
        FftTile<LOGM, THREADS, COLS, false, LIVE>::run(v, t, ll, lds, job.tw, job.chirp_out, pf);
    }

    if (valid) {
        float2* o = job.dst + int64_t(line) * job.ls + int64_t(t) * es;
        const float sc = job.scale;
#pragma unroll
        for (int i = 0; i < LIVE; ++i) {
            const uint32_t pos = uint32_t(t + i * C::TPL);
            if (in_line((int)pos)) {
                float2 y;
                if constexpr (BLUE) {
                    // forward conj(v)*c = conj(v*conj(c)), inverse conj(v*c)/n = conj(v*(c/n))
                    y = cmul_conj(v[i], pf[i]);
                } else {
                    y = cscale(cxor(v[i], sgn), sc);
                }
                if constexpr ((IO & IO_MAG) != 0)
                    io.mag[int64_t(line) * io.mls + int64_t(pos) * io.mel] = cabs_scaled(y, io.mscale);
                else
                    o[i * step] = y;
            }
        }
    }
}

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Give each XCD a contiguous range of tiles so neighbouring column tiles,
// which share cache lines, meet in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    constexpr int NXCD = 8;
    const int xcd = bid % NXCD;
    const int local = bid / NXCD;
    const int q = nblocks / NXCD, r = nblocks % NXCD;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

}  // namespace mdft
