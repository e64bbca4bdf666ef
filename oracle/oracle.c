/*
 * CPU oracle for the 2-D DFT path.  TEST INFRASTRUCTURE ONLY (the checker):
 * loaded by tests, the smoke check and the benchmark's CPU-baseline leg, never
 * by the product library.  Each function cites the reference code it restates.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -pthread).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static const float PI_F = 3.14159265358979323846f;
static const double PI_D = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* 1. Plan math (src/gpu/dft/function.cpp)                                   */
/* ------------------------------------------------------------------------ */

static int is_pow2(unsigned v) { return v != 0 && (v & (v - 1)) == 0; }

static unsigned bit_ceil_u(unsigned v)
{
    unsigned r = 1;
    while (r < v) r <<= 1;
    return r;
}

static unsigned bit_floor_u(unsigned v)
{
    unsigned r = 1;
    if (v == 0) return 0;
    while ((r << 1) != 0 && (r << 1) <= v) r <<= 1;
    return r;
}

static int log2_u(unsigned v)
{
    int r = 0;
    while ((1u << r) < v) ++r;
    return r;
}

/* function.cpp:32-45: n if n is a power of two, else bit_ceil(2n-2);
 * n outside [1, 2^29] is an error. */
int orc_compute_m(int n)
{
    const int max_pow2 = 1 << 30;
    if (n < 1 || n > max_pow2 / 2) return -1;
    if (is_pow2((unsigned)n)) return n;
    return (int)bit_ceil_u((unsigned)(2 * n - 2));
}

/* function.cpp:49-71: h[l] = polar(coef, (inverse ? -pi : pi) * factor),
 * factor = (q mod 2) + r/n where l*l = q*n + r in 64-bit integers. */
void orc_compute_h(int n, int inverse, double coef, double* h)
{
    for (int l = 0; l < n; ++l) {
        const long long dividend = (long long)l * l;
        const long long quotient = dividend / n;
        const long long remainder = dividend - quotient * n;
        const double factor = (double)(quotient & 1) + (double)remainder / n;
        const double theta = (inverse ? -PI_D : PI_D) * factor;
        h[2 * l] = coef * cos(theta);
        h[2 * l + 1] = coef * sin(theta);
    }
}

/* function.cpp:75-92: circulant embedding of length m. */
void orc_compute_h2(int n, int m, const double* h, double* h2)
{
    for (int l = 0; l <= n - 1; ++l) {
        h2[2 * l] = h[2 * l];
        h2[2 * l + 1] = h[2 * l + 1];
    }
    for (int l = n; l <= m - n; ++l) {
        h2[2 * l] = 0.0;
        h2[2 * l + 1] = 0.0;
    }
    for (int l = m - n + 1; l <= m - 1; ++l) {
        h2[2 * l] = h[2 * (m - l)];
        h2[2 * l + 1] = h[2 * (m - l) + 1];
    }
}

/* function.cpp:94-105 with sizeof(complex<float>) = 8. */
unsigned orc_shared_size(unsigned dft_size, unsigned max_shared_memory_bytes)
{
    const unsigned s1 = dft_size > 128u ? dft_size : 128u;
    const unsigned s2 = bit_floor_u(max_shared_memory_bytes / 8u);
    return s1 < s2 ? s1 : s2;
}

/* ------------------------------------------------------------------------ */
/* 2. FP32 restatement of the reference GPU pipeline                         */
/* ------------------------------------------------------------------------ */

/* math.glsl:23-33 */
static unsigned bit_reverse_u(unsigned v, int bits)
{
    unsigned r = 0;
    for (int i = 0; i < bits; ++i) {
        r = (r << 1) | (v & 1u);
        v >>= 1;
    }
    return r;
}

static void cmul_f(float ar, float ai, float br, float bi, float* xr, float* xi)
{
    *xr = ar * br - ai * bi;
    *xi = ar * bi + ai * br;
}

/* dft_bit_reverse.comp:35-56 on one line (the swap i <-> rev(i) for i < rev). */
static void bit_reverse_line32(int n, float* x)
{
    const int bits = log2_u((unsigned)n);
    for (unsigned i = 0; i < (unsigned)n; ++i) {
        const unsigned r = bit_reverse_u(i, bits);
        if (i < r) {
            float t0 = x[2 * i], t1 = x[2 * i + 1];
            x[2 * i] = x[2 * r];
            x[2 * i + 1] = x[2 * r + 1];
            x[2 * r] = t0;
            x[2 * r + 1] = t1;
        }
    }
}

/* dft_fft_shared.comp:90-137 on one tile of nmax elements already in
 * bit-reversed order: radix-2 DIT stages with twiddle
 * (cos(theta*idx), sin(theta*idx)), theta = -+pi halved per stage; the inverse
 * divides by N_MAX on store (:72). */
static void shared_stages32(int nmax, float* x, int inverse)
{
    float two_pi_div_m = inverse ? PI_F : -PI_F;
    for (unsigned m_div_2 = 1; m_div_2 < (unsigned)nmax; m_div_2 <<= 1, two_pi_div_m *= 0.5f) {
        for (unsigned local = 0; local < (unsigned)nmax / 2; ++local) {
            const unsigned idx = local & (m_div_2 - 1);
            const unsigned base = (local - idx) << 1;
            const unsigned it = base + idx;
            const unsigned iu = it + m_div_2;
            const float ang = two_pi_div_m * (float)idx;
            float kr, ki;
            cmul_f(cosf(ang), sinf(ang), x[2 * iu], x[2 * iu + 1], &kr, &ki);
            const float tr = x[2 * it], ti = x[2 * it + 1];
            x[2 * it] = tr + kr;
            x[2 * it + 1] = ti + ki;
            x[2 * iu] = tr - kr;
            x[2 * iu + 1] = ti - ki;
        }
    }
    if (inverse) {
        const float d = (float)nmax;
        for (int i = 0; i < 2 * nmax; ++i) x[i] = x[i] / d;
    }
}

/* dft_fft_global.comp:44-78 for one stage on one line. */
static void global_stage32(int n, float* x, unsigned m_div_2, float two_pi_div_m, int inverse)
{
    for (unsigned i = 0; i < (unsigned)n / 2; ++i) {
        const unsigned idx = i & (m_div_2 - 1);
        const unsigned base = (i - idx) << 1;
        const unsigned it = base + idx;
        const unsigned iu = it + m_div_2;
        const float ang = two_pi_div_m * (float)idx;
        const float c = cosf(ang), s = sinf(ang);
        float kr, ki;
        cmul_f(c, inverse ? s : -s, x[2 * iu], x[2 * iu + 1], &kr, &ki);
        const float tr = x[2 * it], ti = x[2 * it + 1];
        if (!inverse) {
            x[2 * it] = tr + kr;
            x[2 * it + 1] = ti + ki;
            x[2 * iu] = tr - kr;
            x[2 * iu + 1] = ti - ki;
        } else {
            x[2 * it] = (tr + kr) * 0.5f;
            x[2 * it + 1] = (ti + ki) * 0.5f;
            x[2 * iu] = (tr - kr) * 0.5f;
            x[2 * iu + 1] = (ti - ki) * 0.5f;
        }
    }
}

/* fft.cpp:226-262,293-329 for one line: n == 1 is a no-op; n <= shared size
 * runs the LDS kernel with reversed input; otherwise bit reverse, the LDS
 * kernel over tiles of shared size, then log2(n/shared) global stages with
 * theta = pi/m_div_2 (fft.cpp:168-175). */
static void fft_line32(int n, float* x, int inverse, unsigned shared_size)
{
    if (n == 1) return;
    const int only_shared = (unsigned)n <= shared_size;
    const int nmax = only_shared ? n : (int)shared_size;
    bit_reverse_line32(n, x);
    for (int base = 0; base < n; base += nmax) shared_stages32(nmax, x + 2 * base, inverse);
    if (only_shared) return;
    unsigned m_div_2 = shared_size;
    float two_pi_div_m = PI_F / (float)m_div_2;
    for (; m_div_2 < (unsigned)n; two_pi_div_m /= 2, m_div_2 <<= 1) global_stage32(n, x, m_div_2, two_pi_div_m, inverse);
}

void orc_fft32(int count, int n, float* data, int inverse, unsigned shared_size)
{
    for (int c = 0; c < count; ++c) fft_line32(n, data + 2 * (size_t)c * n, inverse, shared_size);
}

/* dft_mul.comp:44-97.  The reference forms l*l in int32 (exact for
 * l < 46341); 64-bit here gives the same values there and stays exact above. */
void orc_mul_data32(float* x, long long l, int n, int inverse)
{
    const long long dividend = l * l;
    const long long quotient = dividend / n;
    const long long remainder = dividend - quotient * n;
    const float factor = (float)remainder / (float)n;
    const float theta = (inverse ? PI_F : -PI_F) * factor;
    float c = cosf(theta), s = sinf(theta);
    if (quotient & 1) {
        c = -c;
        s = -s;
    }
    float yr, yi;
    cmul_f(c, s, x[0], x[1], &yr, &yi);
    x[0] = yr;
    x[1] = yi;
}

/* dft.cpp:57-69,77-111: D = FFT_m(float(h2)) forward, and the inverse table
 * from h(n, inverse, m/n) through the normalised inverse FFT. */
static void diagonal32(int n, int m, int inverse, unsigned lds_bytes, float* d)
{
    double* h = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* h2 = (double*)malloc(sizeof(double) * 2 * (size_t)m);
    orc_compute_h(n, inverse, inverse ? (double)m / n : 1.0, h);
    orc_compute_h2(n, m, h, h2);
    for (int i = 0; i < 2 * m; ++i) d[i] = (float)h2[i];
    fft_line32(m, d, inverse, orc_shared_size((unsigned)m, lds_bytes));
    free(h);
    free(h2);
}

void orc_diagonals32(int n1, int n2, unsigned lds_bytes, float* d1_fwd, float* d1_inv, float* d2_fwd, float* d2_inv)
{
    const int m1 = orc_compute_m(n1), m2 = orc_compute_m(n2);
    diagonal32(n1, m1, 0, lds_bytes, d1_fwd);
    diagonal32(n1, m1, 1, lds_bytes, d1_inv);
    diagonal32(n2, m2, 0, lds_bytes, d2_fwd);
    diagonal32(n2, m2, 1, lds_bytes, d2_inv);
}

/* One line of one dimension of dft.cpp:326-347: to_buffer (chirp + zero
 * pad), FFT(inverse), x D, FFT(!inverse), from_buffer (chirp + crop).  Every
 * reference stage is independent per line, so running a whole line at a time
 * gives the same arithmetic as the stage-by-stage dispatches. */
static void ref_line32(int n, int m, int inverse, const float* d, unsigned shared_size, float* x, size_t stride,
                       float* buf)
{
    for (int i = 0; i < m; ++i) {
        if (i < n) {
            buf[2 * i] = x[2 * (size_t)i * stride];
            buf[2 * i + 1] = x[2 * (size_t)i * stride + 1];
            orc_mul_data32(buf + 2 * i, i, n, inverse);
        } else {
            buf[2 * i] = 0.0f;
            buf[2 * i + 1] = 0.0f;
        }
    }
    fft_line32(m, buf, inverse, shared_size);
    for (int i = 0; i < m; ++i) { /* dft_mul_d.comp:41-54 */
        float yr, yi;
        cmul_f(d[2 * i], d[2 * i + 1], buf[2 * i], buf[2 * i + 1], &yr, &yi);
        buf[2 * i] = yr;
        buf[2 * i + 1] = yi;
    }
    fft_line32(m, buf, !inverse, shared_size);
    for (int i = 0; i < n; ++i) {
        orc_mul_data32(buf + 2 * i, i, n, inverse);
        x[2 * (size_t)i * stride] = buf[2 * i];
        x[2 * (size_t)i * stride + 1] = buf[2 * i + 1];
    }
}

struct ref32_job {
    int n, m, inverse, lines;
    const float* d;
    unsigned shared_size;
    float* data;
    size_t elem_stride, line_stride;
    int first, last;
};

static void* ref32_worker(void* arg)
{
    struct ref32_job* j = (struct ref32_job*)arg;
    float* buf = (float*)malloc(sizeof(float) * 2 * (size_t)j->m);
    for (int l = j->first; l < j->last; ++l)
        ref_line32(j->n, j->m, j->inverse, j->d, j->shared_size, j->data + 2 * j->line_stride * (size_t)l,
                   j->elem_stride, buf);
    free(buf);
    return NULL;
}

static void run_lines32(int n, int m, int inverse, const float* d, unsigned shared_size, float* data, int lines,
                        size_t elem_stride, size_t line_stride, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > lines) nthreads = lines;
    pthread_t th[256];
    struct ref32_job jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].n = n;
        jobs[t].m = m;
        jobs[t].inverse = inverse;
        jobs[t].lines = lines;
        jobs[t].d = d;
        jobs[t].shared_size = shared_size;
        jobs[t].data = data;
        jobs[t].elem_stride = elem_stride;
        jobs[t].line_stride = line_stride;
        jobs[t].first = (int)((long long)lines * t / nthreads);
        jobs[t].last = (int)((long long)lines * (t + 1) / nthreads);
    }
    if (nthreads == 1) {
        ref32_worker(&jobs[0]);
        return;
    }
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, ref32_worker, &jobs[t]);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* dft.cpp:326-347: rows (skipped when w == 1), then columns (skipped when
 * h == 1); diagonals as dft.cpp:77-111. */
int orc_dft2d_ref32(int w, int h, int inverse, float* data, unsigned lds_bytes, int nthreads)
{
    const int m1 = orc_compute_m(w), m2 = orc_compute_m(h);
    if (m1 < 0 || m2 < 0) return -1;
    if (w > 1) {
        float* d = (float*)malloc(sizeof(float) * 2 * (size_t)m1);
        diagonal32(w, m1, inverse, lds_bytes, d);
        run_lines32(w, m1, inverse, d, orc_shared_size((unsigned)m1, lds_bytes), data, h, 1, (size_t)w, nthreads);
        free(d);
    }
    if (h > 1) {
        float* d = (float*)malloc(sizeof(float) * 2 * (size_t)m2);
        diagonal32(h, m2, inverse, lds_bytes, d);
        run_lines32(h, m2, inverse, d, orc_shared_size((unsigned)m2, lds_bytes), data, w, (size_t)w, 1, nthreads);
        free(d);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3. FP64 oracle                                                            */
/* ------------------------------------------------------------------------ */

/* exp(-+2*pi*i*q/n) for an exactly reduced integer q in [0, n). */
static void unit_root64(long long q, long long n, int inverse, double* c, double* s)
{
    /* fold into the first octant-free form: angle = 2*pi*q/n, q < n */
    const double a = 2.0 * PI_D * (double)q / (double)n;
    *c = cos(a);
    *s = inverse ? sin(a) : -sin(a);
}

void orc_dft1d_direct64(int n, int inverse, const double* in, double* out)
{
    double* tw = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (int q = 0; q < n; ++q) unit_root64(q, n, inverse, &tw[2 * q], &tw[2 * q + 1]);
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < n; ++j) {
            const long long q = ((long long)j * k) % n;
            const double c = tw[2 * q], s = tw[2 * q + 1];
            sr += in[2 * j] * c - in[2 * j + 1] * s;
            si += in[2 * j] * s + in[2 * j + 1] * c;
        }
        out[2 * k] = sr;
        out[2 * k + 1] = si;
    }
    if (inverse)
        for (int i = 0; i < 2 * n; ++i) out[i] /= n;
    free(tw);
}

/* In-place iterative radix-2 DIT FFT of a power-of-two length, unnormalised. */
static void fft64_pow2(int n, double* x, int inverse, const double* tw /* n/2 roots of n, forward */)
{
    const int bits = log2_u((unsigned)n);
    for (unsigned i = 0; i < (unsigned)n; ++i) {
        const unsigned r = bit_reverse_u(i, bits);
        if (i < r) {
            double t0 = x[2 * i], t1 = x[2 * i + 1];
            x[2 * i] = x[2 * r];
            x[2 * i + 1] = x[2 * r + 1];
            x[2 * r] = t0;
            x[2 * r + 1] = t1;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int b = 0; b < n; b += len) {
            for (int k = 0; k < half; ++k) {
                const double c = tw[2 * k * step];
                const double s = inverse ? -tw[2 * k * step + 1] : tw[2 * k * step + 1];
                const int it = b + k, iu = b + k + half;
                const double ur = x[2 * iu] * c - x[2 * iu + 1] * s;
                const double ui = x[2 * iu] * s + x[2 * iu + 1] * c;
                const double tr = x[2 * it], ti = x[2 * it + 1];
                x[2 * it] = tr + ur;
                x[2 * it + 1] = ti + ui;
                x[2 * iu] = tr - ur;
                x[2 * iu + 1] = ti - ui;
            }
        }
    }
}

struct line64_plan {
    int n, m, kind; /* kind 0 direct, 1 pow2, 2 bluestein */
    double* tw;     /* m/2 forward roots of m (pow2/bluestein), n roots for direct */
    double* chirp;  /* bluestein: exp(-i*pi*k^2/n), k < n */
    double* dft_b;  /* bluestein: FFT_m of the embedded conjugate chirp */
};

static void line64_plan_init(struct line64_plan* p, int n)
{
    memset(p, 0, sizeof(*p));
    p->n = n;
    if (n <= 64 && !is_pow2((unsigned)n)) {
        p->kind = 0;
        return;
    }
    if (is_pow2((unsigned)n)) {
        p->kind = 1;
        p->m = n;
    } else {
        p->kind = 2;
        p->m = orc_compute_m(n);
    }
    const int m = p->m;
    p->tw = (double*)malloc(sizeof(double) * (size_t)(m > 1 ? m : 2));
    for (int k = 0; k < m / 2; ++k) unit_root64(k, m, 0, &p->tw[2 * k], &p->tw[2 * k + 1]);
    if (p->kind != 2) return;
    p->chirp = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    p->dft_b = (double*)calloc(2 * (size_t)m, sizeof(double));
    for (int k = 0; k < n; ++k) {
        /* k^2 mod 2n in 64-bit integers: exp(-i*pi*k^2/n) = exp(-2*pi*i*(k^2 mod 2n)/(2n)) */
        const long long q = ((long long)k * k) % (2LL * n);
        unit_root64(q, 2LL * n, 0, &p->chirp[2 * k], &p->chirp[2 * k + 1]);
    }
    /* b[l] = conj(chirp[|l|]) on the circulant of length m */
    for (int l = 0; l < n; ++l) {
        p->dft_b[2 * l] = p->chirp[2 * l];
        p->dft_b[2 * l + 1] = -p->chirp[2 * l + 1];
        if (l > 0) {
            p->dft_b[2 * (m - l)] = p->chirp[2 * l];
            p->dft_b[2 * (m - l) + 1] = -p->chirp[2 * l + 1];
        }
    }
    fft64_pow2(m, p->dft_b, 0, p->tw);
}

static void line64_plan_free(struct line64_plan* p)
{
    free(p->tw);
    free(p->chirp);
    free(p->dft_b);
}

/* One line, reference semantics, using scratch of 2*m doubles. */
static void line64_exec(const struct line64_plan* p, int inverse, double* x, size_t stride, double* scratch)
{
    const int n = p->n;
    if (n == 1) return;
    if (p->kind == 0) {
        double* in = scratch;
        double* out = scratch + 2 * n;
        for (int i = 0; i < n; ++i) {
            in[2 * i] = x[2 * i * stride];
            in[2 * i + 1] = x[2 * i * stride + 1];
        }
        orc_dft1d_direct64(n, inverse, in, out);
        for (int i = 0; i < n; ++i) {
            x[2 * i * stride] = out[2 * i];
            x[2 * i * stride + 1] = out[2 * i + 1];
        }
        return;
    }
    const int m = p->m;
    double* a = scratch;
    if (p->kind == 1) {
        for (int i = 0; i < n; ++i) {
            a[2 * i] = x[2 * i * stride];
            a[2 * i + 1] = x[2 * i * stride + 1];
        }
        fft64_pow2(n, a, inverse, p->tw);
        const double sc = inverse ? 1.0 / n : 1.0;
        for (int i = 0; i < n; ++i) {
            x[2 * i * stride] = a[2 * i] * sc;
            x[2 * i * stride + 1] = a[2 * i + 1] * sc;
        }
        return;
    }
    /* Bluestein: X[k] = c[k] * sum_j (c[j] x[j]) conj(c)[k-j] with
     * c = exp(-+i*pi*k^2/n); the inverse conjugates c and the kernel. */
    for (int i = 0; i < m; ++i) {
        if (i < n) {
            const double cr = p->chirp[2 * i], ci = inverse ? -p->chirp[2 * i + 1] : p->chirp[2 * i + 1];
            const double xr = x[2 * i * stride], xi = x[2 * i * stride + 1];
            a[2 * i] = xr * cr - xi * ci;
            a[2 * i + 1] = xr * ci + xi * cr;
        } else {
            a[2 * i] = 0.0;
            a[2 * i + 1] = 0.0;
        }
    }
    fft64_pow2(m, a, 0, p->tw);
    for (int i = 0; i < m; ++i) {
        /* forward kernel conj(c): FFT = dft_b; inverse kernel c: FFT = conj(dft_b)
         * (the circulant kernel is even, so its transform is even) */
        const double br = p->dft_b[2 * i], bi = inverse ? -p->dft_b[2 * i + 1] : p->dft_b[2 * i + 1];
        const double ar = a[2 * i], ai = a[2 * i + 1];
        a[2 * i] = ar * br - ai * bi;
        a[2 * i + 1] = ar * bi + ai * br;
    }
    fft64_pow2(m, a, 1, p->tw);
    const double sc = inverse ? 1.0 / ((double)m * n) : 1.0 / m;
    for (int i = 0; i < n; ++i) {
        const double cr = p->chirp[2 * i], ci = inverse ? -p->chirp[2 * i + 1] : p->chirp[2 * i + 1];
        const double ar = a[2 * i] * sc, ai = a[2 * i + 1] * sc;
        x[2 * i * stride] = ar * cr - ai * ci;
        x[2 * i * stride + 1] = ar * ci + ai * cr;
    }
}

void orc_dft1d_64(int n, int inverse, const double* in, double* out)
{
    struct line64_plan p;
    line64_plan_init(&p, n);
    double* scratch = (double*)malloc(sizeof(double) * 4 * (size_t)(p.m > n ? p.m : n) + 16);
    memcpy(out, in, sizeof(double) * 2 * (size_t)n);
    line64_exec(&p, inverse, out, 1, scratch);
    free(scratch);
    line64_plan_free(&p);
}

struct lines64_job {
    const struct line64_plan* p;
    int inverse;
    double* data;
    size_t elem_stride, line_stride;
    int first, last;
};

static void* lines64_worker(void* arg)
{
    struct lines64_job* j = (struct lines64_job*)arg;
    const int m = j->p->m > j->p->n ? j->p->m : j->p->n;
    double* scratch = (double*)malloc(sizeof(double) * 4 * (size_t)m + 16);
    for (int l = j->first; l < j->last; ++l)
        line64_exec(j->p, j->inverse, j->data + 2 * j->line_stride * (size_t)l, j->elem_stride, scratch);
    free(scratch);
    return NULL;
}

static void run_lines64(int n, int inverse, double* data, int lines, size_t elem_stride, size_t line_stride,
                        int nthreads)
{
    struct line64_plan p;
    line64_plan_init(&p, n);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > lines) nthreads = lines;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    struct lines64_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].p = &p;
        jobs[t].inverse = inverse;
        jobs[t].data = data;
        jobs[t].elem_stride = elem_stride;
        jobs[t].line_stride = line_stride;
        jobs[t].first = (int)((long long)lines * t / nthreads);
        jobs[t].last = (int)((long long)lines * (t + 1) / nthreads);
    }
    if (nthreads == 1)
        lines64_worker(&jobs[0]);
    else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, lines64_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    line64_plan_free(&p);
}

int orc_dft2d_64(int w, int h, int inverse, const double* in, double* out, int nthreads)
{
    if (w < 1 || h < 1) return -1;
    if (out != in) memmove(out, in, sizeof(double) * 2 * (size_t)w * h);
    if (w > 1) run_lines64(w, inverse, out, h, 1, (size_t)w, nthreads);
    if (h > 1) run_lines64(h, inverse, out, w, (size_t)w, 1, nthreads);
    return 0;
}

int orc_dft2d_64_from32(int w, int h, int inverse, const float* in, double* out, int nthreads)
{
    if (w < 1 || h < 1) return -1;
    const size_t count = 2 * (size_t)w * h;
    for (size_t i = 0; i < count; ++i) out[i] = (double)in[i];
    return orc_dft2d_64(w, h, inverse, out, out, nthreads);
}

void orc_error_metrics(size_t count, const float* got, const double* ref, double* out)
{
    double num = 0.0, den = 0.0, maxerr = 0.0, maxref = 0.0, l1 = 0.0, l1g = 0.0;
    for (size_t i = 0; i < count; ++i) {
        const double dr = (double)got[2 * i] - ref[2 * i];
        const double di = (double)got[2 * i + 1] - ref[2 * i + 1];
        const double e2 = dr * dr + di * di;
        const double r2 = ref[2 * i] * ref[2 * i] + ref[2 * i + 1] * ref[2 * i + 1];
        num += e2;
        den += r2;
        const double e = sqrt(e2), r = sqrt(r2);
        if (e > maxerr) maxerr = e;
        if (r > maxref) maxref = r;
        l1 += e;
        l1g += sqrt((double)got[2 * i] * got[2 * i] + (double)got[2 * i + 1] * got[2 * i + 1]);
    }
    out[0] = den > 0 ? sqrt(num / den) : sqrt(num);
    out[1] = maxref > 0 ? maxerr / maxref : maxerr;
    out[2] = maxerr;
    out[3] = (l1 == 0.0) ? 0.0 : l1 / l1g;
}

/* splitmix64 finaliser on (seed, counter). */
static uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void orc_fill(uint64_t seed, int dist, size_t count, float* out)
{
    const uint64_t key = mix64(seed ^ 0x6d646674ULL);
    for (size_t i = 0; i < count; ++i) {
        const uint64_t r = mix64(key + 2 * (uint64_t)i);
        const uint64_t s = mix64(key + 2 * (uint64_t)i + 1);
        if (dist == 0) {
            out[2 * i] = (float)((double)(r >> 40) * (1.0 / 8388608.0) - 1.0);
            out[2 * i + 1] = (float)((double)(s >> 40) * (1.0 / 8388608.0) - 1.0);
        } else {
            const double u1 = ((double)(r >> 11) + 1.0) * (1.0 / 9007199254740992.0);
            const double u2 = (double)(s >> 11) * (1.0 / 9007199254740992.0);
            const double rad = sqrt(-2.0 * log(u1));
            out[2 * i] = (float)(rad * cos(2.0 * PI_D * u2));
            out[2 * i + 1] = (float)(rad * sin(2.0 * PI_D * u2));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Image path: luminance of the source rectangle                             */
/* (src/gpu/dft/code/dft_copy_input.comp:37-54)                              */
/* ------------------------------------------------------------------------ */

void orc_image_luminance(int fmt, const void* src, int64_t pitch, int x, int y, int w, int h, float* lum)
{
    for (int j = 0; j < h; ++j) {
        for (int i = 0; i < w; ++i) {
            const int64_t idx = (int64_t)(y + j) * pitch + (x + i);
            float r, g, b;
            if (fmt == 1) {
                const unsigned char* p = (const unsigned char*)src + 4 * idx;
                r = (float)p[0] / 255.0f;
                g = (float)p[1] / 255.0f;
                b = (float)p[2] / 255.0f;
            } else {
                const float* p = (const float*)src + 4 * idx;
                r = p[0];
                g = p[1];
                b = p[2];
            }
            lum[(int64_t)j * w + i] = fmaf(0.0722f, b, fmaf(0.7152f, g, 0.2126f * r));
        }
    }
}
