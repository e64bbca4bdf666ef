// Device-side building blocks of the 2-D DFT line passes (gfx950 / CDNA4).
//
// One "axis pass" transforms `count` lines of logical length n, laid out with
// element stride es and line stride ls inside a row-major complex<float>
// image (rows: es = 1, ls = w; columns: es = w, ls = 1).  A line of length n
// is transformed through an FFT of length M: directly when n is a power of
// two or 3 * 2^k in [192, 6144] (M = n), otherwise by Bluestein's chirp-z
// method with M >= 2n - 1: the reference's padded length bit_ceil(2n - 2)
// (src/gpu/dft/function.cpp:32-45), or 3 * 2^k when that is shorter, and the
// chirp tables of function.cpp:49-92 generated in FP64 on the device.
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
// A line's FFT has a "shape" S:
//  * S in [1, MAX_LOGM]: M = 2^S, radix-16 passes (16 points per thread) and
//    one remainder pass;
//  * S = F3_BASE + k: M = 3 * 2^k (Bluestein convolutions, and lines of
//    exactly 3 * 2^k points transformed directly), 12 points per
//    thread: radix-4 passes, one radix-2 pass when k is odd, and a final
//    radix-12 pass (a 3 x 4 prime-factor butterfly).
// The shape is a function of the line length and axis, so a line runs the
// same schedule in every kernel.
constexpr int MAX_LOGM = 14;
// waves per SIMD the tile kernels are compiled for (at most 128 VGPRs)
constexpr int TILE_WAVES = 4;
constexpr int F3_BASE = 32;
constexpr int F3_MIN_K = 6;   // M = 192: the radix-12 pass follows at least two radix-4 passes
constexpr int F3_MAX_K = 11;  // M = 6144: 512 threads per line
__host__ __device__ constexpr bool is_f3(int s) { return s >= F3_BASE; }
__host__ __device__ constexpr int f3_k(int s) { return s - F3_BASE; }
// bits per pass of the full-radix passes
__host__ __device__ constexpr int fft_bits(int s) { return is_f3(s) ? 2 : 4; }
__host__ __device__ constexpr int m_of(int s) { return is_f3(s) ? (3 << f3_k(s)) : (1 << s); }
__host__ __device__ constexpr int ept_of(int s)
{
    return is_f3(s) ? 12 : s < fft_bits(s) ? (1 << s) : (1 << fft_bits(s));
}
__host__ __device__ constexpr int npass_of(int s)
{
    return is_f3(s) ? (f3_k(s) - 2) / 2 + (f3_k(s) - 2) % 2 + 1 : (s + fft_bits(s) - 1) / fft_bits(s);
}
// radix of pass p (1 when p is past the last pass)
__host__ __device__ constexpr int radix_of(int s, int p)
{
    if (p >= npass_of(s)) return 1;
    if (is_f3(s)) {
        const int a = f3_k(s) - 2;
        return p < a / 2 ? 4 : (p == npass_of(s) - 1 ? 12 : 2);
    }
    const int b = fft_bits(s);
    return 1 << ((s - b * p) >= b ? b : (s - b * p));
}
// Stockham stride before pass p: the product of the earlier radices
__host__ __device__ constexpr int ns_of(int s, int p)
{
    int ns = 1;
    for (int q = 0; q < p; ++q) ns *= radix_of(s, q);
    return ns;
}

// Twiddle rows of pass p >= 1 (stride Ns = ns_of(s, p), radix R): row k
// (k < Ns) holds TW_ROW = 4 entries w^1, w^2, w^3, w^4 of
// w = exp(-2*pi*i*k/(Ns*R)) (32 bytes, four rows per cache line); the other
// powers are products of these (expand_twiddles).
constexpr int TW_ROW = 4;
__host__ __device__ constexpr int64_t tw_pass_len(int s, int p)
{
    return (p >= 1 && p < npass_of(s) ? int64_t(ns_of(s, p)) : 0) * TW_ROW;
}
__host__ __device__ constexpr int64_t tw_pass_off(int s, int p)
{
    int64_t off = 0;
    for (int q = 1; q < p; ++q) off += tw_pass_len(s, q);
    return off;
}
__host__ __device__ constexpr int64_t tw_len(int s)
{
    int64_t t = 0;
    for (int p = 1; p < npass_of(s); ++p) t += tw_pass_len(s, p);
    return t;
}
// Every shape has a block of one table: powers of two 1..MAX_LOGM, then
// 3 * 2^k for k in [F3_MIN_K, F3_MAX_K].
__host__ __device__ constexpr int64_t tw_off(int s)
{
    int64_t t = 0;
    const int plast = is_f3(s) ? MAX_LOGM + 1 : s;
    for (int l = 1; l < plast; ++l) t += tw_len(l);
    if (is_f3(s))
        for (int k = F3_MIN_K; k < f3_k(s); ++k) t += tw_len(F3_BASE + k);
    return t;
}
__host__ __device__ constexpr int64_t tw_total() { return tw_off(F3_BASE + F3_MAX_K + 1); }

// LDS layout of the exchange written by pass P: line position p is stored at
// word p + c * (p >> sh) of the line's slice, with (c, sh) chosen per FFT
// shape, workgroup size, axis and pass by gen_lds_maps.py (lds_maps.inc) so
// that (a) offsets stay linear, f(t + i*TPL) = f(t) + f(i*TPL) for the next
// pass's reads and f(base + r*NS) = f(base) + r*f(NS) for this pass's writes,
// making every LDS access a base register plus an immediate, and (b) no
// write or read instruction has a bank conflict (32 banks, 32 lanes per cycle).
struct LdsMap {
    int shape, threads, cols, pass, c, sh;
};
constexpr LdsMap kLdsMaps[] = {
#include "lds_maps.inc"
};
__host__ __device__ constexpr LdsMap lds_map(int shape, int threads, bool cols, int pass)
{
    for (const LdsMap& e : kLdsMaps)
        if (e.shape == shape && e.threads == threads && e.cols == (cols ? 1 : 0) && e.pass == pass) return e;
    return LdsMap{shape, threads, cols ? 1 : 0, pass, -1, 0};  // missing: rejected by a static_assert
}
template <int S, int THREADS, bool COLS, int P>
__host__ __device__ constexpr int lds_pos(int p)
{
    constexpr LdsMap m = lds_map(S, THREADS, COLS, P);
    static_assert(m.c >= 0, "no LDS layout generated for this shape, workgroup size and pass");
    return p + m.c * (p >> m.sh);
}

template <int S, int THREADS>
struct LineCfg {
    static constexpr int M = m_of(S);
    static constexpr int EPT = ept_of(S);
    static constexpr int TPL = M / EPT;          // threads per line
    static constexpr int LPW = THREADS / TPL;    // lines per tile
    static constexpr int NPASS = npass_of(S);
    static_assert(TPL <= THREADS, "line does not fit the workgroup");
    static_assert(LPW * TPL == THREADS, "workgroup must hold whole lines");
    // words of one line's slice: the largest padded position over the exchanges
    template <bool COLS, int P = 0>
    static constexpr int padm()
    {
        if constexpr (P + 1 >= NPASS) {
            return M;
        } else {
            const int a = lds_pos<S, THREADS, COLS, P>(M - 1) + 1;
            const int b = padm<COLS, P + 1>();
            return a > b ? a : b;
        }
    }
    // Line stride in LDS words.  Lanes of one wave cover q consecutive
    // positions of 32/q lines (rows: q = TPL; columns, lines fastest:
    // q = 32/LPW); a stride of q times an odd number puts them on distinct
    // banks (the same rule gen_lds_maps.py assumes).
    template <bool COLS>
    static constexpr int lstride()
    {
        const int q = COLS ? (LPW >= 32 ? 1 : 32 / LPW) : (TPL >= 32 ? 0 : TPL);
        const int pm = padm<COLS>();
        if (q == 0) return pm;
        int s = pm;
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

// forward 3-point DFT in place, W = exp(-2*pi*i/3): X1 = t2 - i*t3,
// X2 = t2 + i*t3 with t2 = x0 - (x1 + x2)/2, t3 = sin(pi/3) * (x1 - x2)
__device__ __forceinline__ void dft3(float2& x0, float2& x1, float2& x2)
{
    const float h = 0.86602540378443864676f;
    const float2 t1 = cadd(x1, x2);
    const float2 t2 = make_float2(__builtin_fmaf(-0.5f, t1.x, x0.x), __builtin_fmaf(-0.5f, t1.y, x0.y));
    const float2 t3 = make_float2(h * (x1.x - x2.x), h * (x1.y - x2.y));
    x0 = cadd(x0, t1);
    x1 = make_float2(t2.x + t3.y, t2.y - t3.x);
    x2 = make_float2(t2.x - t3.y, t2.y + t3.x);
}

// ZHI: inputs u[R/2 + 1 ..] are zero (the zero-padded half of a Bluestein
// line in its first pass).  R = 12 is the prime-factor (Good-Thomas) 3 x 4
// DFT: no twiddles; input n = (4*n1 + 3*n2) mod 12, output k = (4*k1 + 9*k2)
// mod 12 (k1 = k mod 3, k2 = k mod 4).
template <int R, bool ZHI = false>
__device__ __forceinline__ void dft_fwd(float2* u)
{
    if constexpr (R == 12) {
        static_assert(!ZHI, "no zero-input pruning for radix 12 (never a first pass)");
        float2 a0[4] = {u[0], u[3], u[6], u[9]};
        float2 a1[4] = {u[4], u[7], u[10], u[1]};
        float2 a2[4] = {u[8], u[11], u[2], u[5]};
        dft_fwd<4>(a0);
        dft_fwd<4>(a1);
        dft_fwd<4>(a2);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            dft3(a0[k2], a1[k2], a2[k2]);
            u[(9 * k2) % 12] = a0[k2];
            u[(4 + 9 * k2) % 12] = a1[k2];
            u[(8 + 9 * k2) % 12] = a2[k2];
        }
    } else {
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
    const float2* tw;      // twiddle block of the line's FFT shape
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
template <int S>
__host__ __device__ constexpr bool tw_prefetch()
{
    return is_f3(S) || S <= 13;
}

// float4s of a row a radix-R butterfly reads: (w^1, w^2) and (w^3, w^4)
template <int R>
__host__ __device__ constexpr int tw_quads()
{
    return R <= 2 ? 1 : 2;
}

template <int S, int THREADS, int P>
struct TwLoad {
    using C = LineCfg<S, THREADS>;
    static constexpr int R = radix_of(S, P);
    static constexpr int Q = tw_quads<R>();
    static constexpr int NB = C::EPT / R;
    static constexpr int NS = ns_of(S, P);
    static_assert(NB * Q <= C::EPT / 2, "twiddle registers");
    __device__ __forceinline__ static void run(float4 (&w)[C::EPT / 2], int t, const float2* tw)
    {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int k = (t + b * C::TPL) & (NS - 1);
            const float4* row = reinterpret_cast<const float4*>(tw + tw_pass_off(S, P) + int64_t(k) * TW_ROW);
#pragma unroll
            for (int r2 = 0; r2 < Q; ++r2) w[b * Q + r2] = row[r2];
        }
    }
};

// the R twiddles of one butterfly (wt[0] unused) from its row: w^1..w^4
// loaded, w^(4a+b) = w^(4a) * w^b, w^(4a) = w^(4a-4) * w^4 (at most four
// roundings from the FP64-generated entries)
template <int R, int r>
__device__ __forceinline__ void expand_products(float2 (&wt)[R])
{
    if constexpr (r < R) {
        if constexpr (r % 4 == 0)
            wt[r] = cmul(wt[r - 4], wt[4]);
        else
            wt[r] = cmul(wt[r - r % 4], wt[r % 4]);
        expand_products<R, r + 1>(wt);
    }
}

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
        expand_products<R, 5>(wt);
    }
}

// ZIN (first pass only): slots EPT/2 + 1 .. EPT of every thread are zero.
// PF: the first PF slots of a table (pfs, indexed by line position) are
// loaded into pf one pass before the last, so their latency overlaps the
// last exchange and butterflies; slots PF .. PFL-1 at the start of the last
// pass, after its exchange (with one pass, all PFL at its start).
template <int S, int THREADS, bool COLS, int P, bool ZIN = false, int PF = 0, int PFL = 0>
struct FftPass {
    using C = LineCfg<S, THREADS>;
    static constexpr bool WAVE = !COLS && C::TPL <= 64;
    static constexpr int R = radix_of(S, P);
    static constexpr int NB = C::EPT / R;
    static constexpr int NS = ns_of(S, P);
    static constexpr bool LAST = (P == C::NPASS - 1);
    static constexpr int LST = C::template lstride<COLS>();
    static_assert(NB * R == C::EPT, "the radix divides the points per thread");

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
            // layout written by the previous pass; the offset of slot i is a
            // compile-time constant (an LDS immediate)
            const int q0 = lds_pos<S, THREADS, COLS, P - 1>(t);
#pragma unroll
            for (int i = 0; i < C::EPT; ++i) {
                const int q = q0 + lds_pos<S, THREADS, COLS, P - 1>(i * C::TPL);
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
                const float4* row = reinterpret_cast<const float4*>(tw + tw_pass_off(S, P) + int64_t(k) * TW_ROW);
#pragma unroll
                for (int r2 = 0; r2 < Q; ++r2) q[r2] = tw_prefetch<S>() ? w[b * Q + r2] : row[r2];
                float2 wt[R];
                expand_twiddles<R>(wt, q);
#pragma unroll
                for (int r = 1; r < R; ++r) u[r] = cmul(u[r], wt[r]);
            }
            // ZIN: slot b + NB*r holds a zero for r > R/2 (b + NB*(R/2 + 1) >= EPT/2 + 1)
            static_assert(!ZIN || P == 0, "zero inputs only in the first pass");
            dft_fwd<R, ZIN>(u);
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + NB * r] = u[r];
        }
        if constexpr (!LAST) {
            if constexpr (tw_prefetch<S>()) TwLoad<S, THREADS, P + 1>::run(w, t, tw);
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
                // the layout keeps the R outputs' offsets linear: each step adds f(NS)
                constexpr int step = lds_pos<S, THREADS, COLS, P>(NS);
                const int q0 = lds_pos<S, THREADS, COLS, P>(base);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int q = q0 + r * step;
                    re[q] = v[b + NB * r].x;
                    im[q] = v[b + NB * r].y;
                }
            }
            FftPass<S, THREADS, COLS, P + 1, false, PF, PFL>::run(v, t, re, im, tw, w, pfs, pf);
        }
    }
};

// ZIN: the upper half of the line is zero apart from position M/2, so slots
// EPT/2 + 1 .. EPT of every thread hold zeros (the caller set them to
// constant zero).  PF > 0: pf[i] = pfs[t + i*TPL] for i < max(PF, PFL) on
// return, loaded during the FFT (pfs must be readable at every such
// position).
template <int S, int THREADS, bool COLS, bool ZIN = false, int PF = 0, int PFL = 0>
struct FftTile {
    using C = LineCfg<S, THREADS>;
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
        if constexpr (tw_prefetch<S>() && C::NPASS > 1) TwLoad<S, THREADS, 1>::run(w, t, tw);
        FftPass<S, THREADS, COLS, 0, ZIN, PF, PFL>::run(v, t, re, im, tw, w, pfs, pf);
    }
};

// ---------------------------------------------------------------------------
// one tile of one axis pass: load (+chirp), FFT, [x D, FFT], store (+chirp)
// ---------------------------------------------------------------------------
// IO (image kernels only): IO_LUM loads luminance from io.img instead of
// complex values from job.src; IO_MAG stores |X| * io.mscale to io.mag
// instead of complex values to job.dst.
template <int S, int THREADS, bool BLUE, bool COLS, int IO = 0>
__device__ __forceinline__ void line_tile(const AxisJob& job, int tile, float* lds, int tid, ImageIO io = ImageIO{})
{
    using C = LineCfg<S, THREADS>;
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
    // a direct line (2^k or 3 * 2^k points) fills the FFT (n == M); only Bluestein pads
    auto in_line = [&](int pos) { return !BLUE || pos < n; };
    // Bluestein: n <= M/2 + 1, so slot i holds a position of the line only
    // for i <= EPT/2 (slot EPT/2 only at t = 0 when n = M/2 + 1; for 3 * 2^k
    // lengths M >= 2n - 1, so i < EPT/2); the other slots are zero on input
    // and not stored on output.
    constexpr int LIVE = BLUE ? (is_f3(S) ? C::EPT / 2 : C::EPT / 2 + 1) : C::EPT;
    // the first FFT skips the zero half in the butterflies of its first pass
    constexpr bool ZIN = BLUE;

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
                // forward x*c, inverse x*conj(c) (direction tables); direct lines: conj for the inverse
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
    FftTile<S, THREADS, COLS, ZIN, PFD, PFDL>::run(v, t, ll, lds, job.tw, job.dtab, pf);

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
        FftTile<S, THREADS, COLS, false, LIVE>::run(v, t, ll, lds, job.tw, job.chirp_out, pf);
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
