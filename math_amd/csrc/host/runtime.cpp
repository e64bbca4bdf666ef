// Per-device context: twiddle tables, the cache of per-length line plans
// (chirp and diagonal tables generated in FP64 on the device), and the
// enqueueing of axis passes.
#include "runtime.hpp"

#include "../../../include/mdft.h"

#include <cstdlib>
#include <string>
#include <vector>

namespace mdft {

constexpr int64_t LineMaxTile = int64_t(1) << MAX_LOGM;

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(MDFT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

bool is_pow2(uint64_t v) { return v != 0 && (v & (v - 1)) == 0; }

int ilog2(uint64_t v)
{
    int r = 0;
    while ((uint64_t(1) << r) < v) ++r;
    return r;
}

int compute_m(long long n)
{
    // src/gpu/dft/function.cpp:32-45
    const long long max_pow2 = 1LL << 30;
    if (n < 1 || n > max_pow2 / 2) throw Error(MDFT_ERR_SIZE, "Error size " + std::to_string(n) + " for compute m");
    if (is_pow2((uint64_t)n)) return (int)n;
    uint64_t m = 1;
    while (m < (uint64_t)(2 * n - 2)) m <<= 1;
    return (int)m;
}

int three_factor_min_logm()
{
    static const int v = [] {
        const char* e = std::getenv("MDFT_LARGE_SPLIT3_LOGM");
        const int x = e ? std::atoi(e) : 0;
        return x >= 15 && x <= 29 ? x : 29;
    }();
    return v;
}

void large_split(int logm, int* l1, int* l2, int* l3)
{
    if (logm >= three_factor_min_logm()) {
        *l3 = logm / 3;
        *l2 = (logm - *l3) / 2;
        *l1 = logm - *l3 - *l2;
    } else {
        *l1 = logm - logm / 2;
        *l2 = logm / 2;
        *l3 = 0;
    }
}

uint64_t large_scratch_elems(long long m, long long count)
{
    return m > LineMaxTile ? (uint64_t)count * (uint64_t)m : 0;
}

int line_shape(long long n, bool cols)
{
    const int m = compute_m(n);
    if (is_pow2((uint64_t)n) || m > (1 << MAX_LOGM)) return ilog2((uint64_t)m);
    // n = 3 * 2^k itself: a direct FFT of that shape, no Bluestein
    if (n % 3 == 0 && is_pow2((uint64_t)(n / 3)) && n >= (3LL << F3_MIN_K) && n <= (3LL << F3_MAX_K))
        return F3_BASE + ilog2((uint64_t)(n / 3));
    long long m3 = 3LL << F3_MIN_K;
    while (m3 < 2 * n - 1) m3 <<= 1;
    const long long min3 = cols ? (3LL << F3_MIN_K) : (3LL << F3_ROW_MIN_K);
    if (m3 < m && m3 >= min3 && m3 <= (3LL << F3_MAX_K)) return F3_BASE + ilog2((uint64_t)(m3 / 3));
    return ilog2((uint64_t)m);
}

LinePlan::~LinePlan()
{
    if (chirp) (void)hipFree(chirp);
    if (dtab) (void)hipFree(dtab);
    if (dir) (void)hipFree(dir);
    if (tw_lo) (void)hipFree(tw_lo);
    if (tw_hi) (void)hipFree(tw_hi);
}

DeviceGuard::DeviceGuard(int device)
{
    hip_check(hipGetDevice(&prev_), "hipGetDevice");
    if (prev_ != device) hip_check(hipSetDevice(device), "hipSetDevice");
}

DeviceGuard::~DeviceGuard()
{
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev_) (void)hipSetDevice(prev_);
}

bool ScratchOrder::capturing(hipStream_t s) const
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(s, &st), "hipStreamIsCapturing");
    return st != hipStreamCaptureStatusNone;
}

void ScratchOrder::before(hipStream_t s)
{
    if (!pending_ || s == last_ || capturing(s)) return;
    hip_check(hipStreamWaitEvent(s, ev_, 0), "hipStreamWaitEvent");
}

void ScratchOrder::after(hipStream_t s)
{
    if (capturing(s)) return;
    if (!ev_) hip_check(hipEventCreateWithFlags(&ev_, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventRecord(ev_, s), "hipEventRecord");
    last_ = s;
    pending_ = true;
}

void ScratchOrder::release()
{
    if (ev_) (void)hipEventDestroy(ev_);
    ev_ = nullptr;
    pending_ = false;
}

namespace {
thread_local int t_device = -1;
std::mutex g_ctx_mu;
std::unordered_map<int, Context*> g_ctx;
}  // namespace

int Context::current_device() { return t_device >= 0 ? t_device : 0; }

Context* Context::find(int device)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(device);
    return it == g_ctx.end() ? nullptr : it->second;
}

void Context::set_thread_device(int device) { t_device = device; }

Context& Context::get(int device)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(device);
    if (it != g_ctx.end()) return *it->second;
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count)
        throw Error(MDFT_ERR_INVALID, "No HIP device " + std::to_string(device) + " (" + std::to_string(count) +
                                          " present)");
    // contexts live until the process ends; device memory is released by the runtime then
    Context* c = new Context(device);
    g_ctx[device] = c;
    return *c;
}

Context::Context(int device) : device_(device)
{
    DeviceGuard g(device);
    hip_check(hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, device), "CU count");
    hip_check(hipStreamCreateWithFlags(&gen_stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipMalloc(&twiddles_, sizeof(float2) * (size_t)tw_total()), "hipMalloc twiddles");
    hip_check(gen_twiddles(twiddles_, gen_stream_), "twiddle generation");
    hip_check(hipStreamSynchronize(gen_stream_), "twiddle generation");
    // FP64 table scratch for every tile-resident Bluestein length (2n + m doubles)
    scratch64_doubles_ = 2 * (size_t(LineMaxTile) / 2 + 1) + size_t(LineMaxTile);
    hip_check(hipMalloc(&scratch64_, sizeof(double) * scratch64_doubles_), "hipMalloc scratch");
    // load every kernel file's code now, not in a caller's first transform
    uint32_t* word = reinterpret_cast<uint32_t*>(scratch64_);
    hip_check(warm_axis(word, gen_stream_), "kernel load");
    hip_check(warm_image(word, gen_stream_), "kernel load");
    hip_check(warm_large(word, gen_stream_), "kernel load");
    hip_check(warm_batch(word, gen_stream_), "kernel load");
    hip_check(hipStreamSynchronize(gen_stream_), "kernel load");
    noalloc_setup(device);
    // default tile shapes: wide tiles for columns of 1024 points and more
    for (int i = 0; i < NSHAPES; ++i) {
        const int shape = i <= MAX_LOGM ? i : F3_BASE + F3_MIN_K + (i - MAX_LOGM - 1);
        for (int b = 0; b < 2; ++b) variant_[i][b][1] = (m_of(shape) >= 1024) ? 1 : 0;
    }
}

static int64_t line_key(int n, int shape) { return (int64_t(shape) << 32) | uint32_t(n); }

std::shared_ptr<const LinePlan> Context::line_plan(int n, bool cols)
{
    const int m = compute_m(n);  // validates n
    const int shape = line_shape(n, cols);
    std::lock_guard<std::mutex> lk(mu_);
    auto it = lines_.find(line_key(n, shape));
    if (it != lines_.end()) {
        it->second.last_use = ++use_clock_;
        return it->second.plan;
    }

    auto lp = std::make_shared<LinePlan>();
    lp->n = n;
    lp->shape = shape;
    lp->padded = m;
    lp->m = is_f3(lp->shape) ? m_of(lp->shape) : m;
    lp->logm = is_f3(lp->shape) ? 0 : ilog2((uint64_t)m);
    // powers of two and n = 3 * 2^k run a direct FFT of length n
    lp->blue = lp->m != n;
    if (lp->logm > LARGE_MAX_LOGM)
        throw Error(MDFT_ERR_SIZE, "Error size " + std::to_string(n) + ": padded length " + std::to_string(m) +
                                       " exceeds the longest supported line (2^" + std::to_string(LARGE_MAX_LOGM) +
                                       ")");
    DeviceGuard g(device_);
    if (lp->large()) {
        large_split(lp->logm, &lp->logm1, &lp->logm2, &lp->logm3);
        const int64_t nhi = int64_t(m) >> LARGE_LO_BITS;
        hip_check(hipMalloc(&lp->tw_lo, sizeof(float2) * LARGE_LO), "hipMalloc twiddles");
        hip_check(hipMalloc(&lp->tw_hi, sizeof(float2) * nhi), "hipMalloc twiddles");
        lp->bytes += sizeof(float2) * (uint64_t)(LARGE_LO + nhi);
        hip_check(gen_large_twiddles(lp->tw_lo, lp->tw_hi, m, gen_stream_), "twiddle generation");
    }
    if (lp->blue) {
        struct DevFree {
            void operator()(void* p) const { (void)hipFree(p); }
        };
        // the FP64 scratch is sized at context creation for every tile-resident
        // length; a long line's FP64 chirp (2n doubles) lives only while its
        // tables are generated
        std::unique_ptr<double, DevFree> chirp64_long;
        if (lp->large()) {
            double* raw64 = nullptr;
            hip_check(hipMalloc(&raw64, sizeof(double) * 2 * (size_t)n), "hipMalloc FP64 chirp");
            chirp64_long.reset(raw64);
        }
        const size_t need = 2 * (size_t)n + (size_t)lp->m;
        if (!lp->large() && need > scratch64_doubles_) {
            hip_check(hipStreamSynchronize(gen_stream_), "table generation");
            if (scratch64_) hip_check(hipFree(scratch64_), "hipFree");
            scratch64_ = nullptr;
            scratch64_doubles_ = 0;
            hip_check(hipMalloc(&scratch64_, sizeof(double) * need), "hipMalloc scratch");
            scratch64_doubles_ = need;
        }
        const int cm = lp->m;  // convolution length
        hip_check(hipMalloc(&lp->chirp, sizeof(float2) * (size_t)n), "hipMalloc chirp");
        hip_check(hipMalloc(&lp->dtab, sizeof(float2) * (size_t)cm), "hipMalloc dtab");
        lp->bytes += sizeof(float2) * ((size_t)n + (size_t)cm);
        double2* chirp64 = reinterpret_cast<double2*>(lp->large() ? chirp64_long.get() : scratch64_);
        if (!lp->large()) {
            hip_check(hipMalloc(&lp->dir, sizeof(float2) * (size_t)dir_tables_len(n, cm)), "hipMalloc tables");
            lp->bytes += sizeof(float2) * (size_t)dir_tables_len(n, cm);
        }
        hip_check(gen_chirp(chirp64, lp->chirp, lp->dir, n, cm, gen_stream_), "chirp generation");
        if (!lp->large()) {
            // D = FFT_M(h2)/M by a direct FP64 sum with exact index reduction
            hip_check(gen_dtab(lp->dtab, chirp64, scratch64_ + 2 * (size_t)n, n, cm, gen_stream_),
                      "diagonal generation");
            hip_check(gen_dir_diagonal(lp->dir, lp->dtab, n, cm, gen_stream_), "diagonal tables");
        } else {
            // long lines: D = FFT_M(h2)/M by this library's four-step FFT on the
            // FP32-rounded embedding (the reference computes its D with an FP32 FFT
            // as well, dft.cpp:57-69)
            float2* raw = nullptr;
            hip_check(hipMalloc(&raw, sizeof(float2) * 2 * (size_t)m), "hipMalloc diagonal scratch");
            std::unique_ptr<float2, DevFree> tmp(raw);
            hip_check(gen_h2_f32(tmp.get(), chirp64, n, m, gen_stream_), "embedding");
            LargeFft f;
            f.m = m;
            f.logm1 = lp->logm1;
            f.logm2 = lp->logm2;
            f.logm3 = lp->logm3;
            f.tw_lo = lp->tw_lo;
            f.tw_hi = lp->tw_hi;
            enqueue_large_fft(f, tmp.get(), lp->dtab, 1, m, 1, 1.0f / (float)m, tmp.get() + m, gen_stream_);
            hip_check(hipStreamSynchronize(gen_stream_), "diagonal generation");
        }
    }
    hip_check(hipStreamSynchronize(gen_stream_), "table generation");
    lines_[line_key(n, shape)] = CacheEntry{lp, ++use_clock_};
    trim_locked(cache_limit_);
    return lp;
}

std::shared_ptr<const LinePlan> Context::find_line_plan(int n, bool cols)
{
    if (n < 1 || n > (1 << 29)) return nullptr;
    const int shape = line_shape(n, cols);
    std::lock_guard<std::mutex> lk(mu_);
    auto it = lines_.find(line_key(n, shape));
    if (it == lines_.end()) return nullptr;
    it->second.last_use = ++use_clock_;
    return it->second.plan;
}

// Free the least recently requested lengths that no plan, batch or server
// holds until the cached bytes are within limit (hipFree waits for the device,
// so no launch that still reads a table is affected).
void Context::trim_locked(uint64_t limit)
{
    uint64_t total = 0;
    for (auto& kv : lines_) total += kv.second.plan->bytes;
    while (total > limit) {
        auto victim = lines_.end();
        for (auto it = lines_.begin(); it != lines_.end(); ++it)
            if (it->second.plan.use_count() == 1 &&
                (victim == lines_.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == lines_.end()) break;  // everything left is in use
        total -= victim->second.plan->bytes;
        lines_.erase(victim);
    }
}

uint64_t Context::cached_bytes()
{
    std::lock_guard<std::mutex> lk(mu_);
    uint64_t s = sizeof(float2) * (uint64_t)tw_total();
    for (auto& kv : lines_) s += kv.second.plan->bytes;
    return s;
}

uint32_t Context::cached_lengths()
{
    std::lock_guard<std::mutex> lk(mu_);
    return (uint32_t)lines_.size();
}

uint64_t Context::cache_limit()
{
    std::lock_guard<std::mutex> lk(mu_);
    return cache_limit_;
}

void Context::set_cache_limit(uint64_t bytes)
{
    std::lock_guard<std::mutex> lk(mu_);
    cache_limit_ = bytes;
    DeviceGuard g(device_);
    trim_locked(bytes);
}

AxisJob Context::make_job(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls, int count,
                          bool inverse) const
{
    AxisJob j;
    j.src = src;
    j.dst = dst;
    j.es = es;
    j.ls = ls;
    j.n = lp.n;
    j.count = count;
    j.inverse = inverse ? 1 : 0;
    j.scale = inverse ? 1.0f / (float)lp.n : 1.0f;
    // Bluestein tables of this direction (tile-resident lengths)
    const int64_t n = lp.n, m = lp.m;
    j.chirp = lp.dir ? lp.dir + (inverse ? n : 0) : nullptr;
    j.dtab = lp.dir ? lp.dir + 2 * n + (inverse ? m : 0) : nullptr;
    j.chirp_out = lp.dir ? lp.dir + 2 * n + 2 * m + (inverse ? n : 0) : nullptr;
    j.tw = twiddles(lp.shape);
    return j;
}

// Step of a long-line FFT: sub-FFTs of length 2^sublogm.
static void large_step(int sublogm, const float2* tw, LargeJob job, int cu_count, hipStream_t stream)
{
    job.tw = tw;
    hip_check(launch_large_step(sublogm, job, cu_count, stream), "long-line step launch");
}

// a step with no twiddle and two-level addressing
static LargeJob large_job(const float2* tw_lo, const float2* tw_hi, bool inverse, int64_t n)
{
    LargeJob j{};
    j.tmod = 1, j.tdiv = 1, j.tmul = 1;
    j.scale = 1.0f;
    j.inverse = inverse ? 1 : 0;
    j.n = n;
    j.tw_lo = tw_lo, j.tw_hi = tw_hi;
    return j;
}

void Context::enqueue_large_a1a2(const LargeFft& f, const float2* src, int64_t es, int64_t ls, int64_t count,
                                 const LargeEnds& e, float2* T, hipStream_t stream) const
{
    const int64_t M = f.m, M1 = int64_t(1) << f.logm1, M2 = int64_t(1) << f.logm2, M3 = int64_t(1) << f.logm3;
    // A1: per (line, j = n1 + M1*n2) the M3-point FFT over n3 (input stride
    // M1*M2), times w_Mr^(n2*k3) = w_M^(M1*n2*k3), into T[line*M + j*M3 + k3]
    LargeJob a = large_job(f.tw_lo, f.tw_hi, e.inverse, e.n);
    a.src = src, a.dst = T;
    a.sub_in = M1 * M2, a.ls_in = ls, a.ls2_in = es, a.es_in = M1 * M2 * es;
    a.sub_out = M1 * M2, a.ls_out = M, a.ls2_out = M3, a.es_out = 1;
    a.count = count * M1 * M2;
    a.load_mode = e.load_mode, a.store_mode = LARGE_STORE_TWIDDLE;
    a.tmod = M1 * M2, a.tdiv = M1, a.tmul = M1;
    a.m1 = M1 * M2;  // the chirp load's element index: j + M1*M2*n3
    a.chirp = e.chirp;
    large_step(f.logm3, twiddles(f.logm3), a, cu_count_, stream);
    // A2, in place: per (line, q = n1*M3 + k3) the M2-point FFT over n2 (stride M1*M3)
    LargeJob b = large_job(f.tw_lo, f.tw_hi, e.inverse, e.n);
    b.src = T, b.dst = T;
    b.sub_in = M1 * M3, b.ls_in = M, b.ls2_in = 1, b.es_in = M1 * M3;
    b.sub_out = M1 * M3, b.ls_out = M, b.ls2_out = 1, b.es_out = M1 * M3;
    b.count = count * M1 * M3;
    b.load_mode = LARGE_LOAD_PLAIN, b.store_mode = LARGE_STORE_PLAIN;
    large_step(f.logm2, twiddles(f.logm2), b, cu_count_, stream);
}

void Context::enqueue_large_fft3(const LargeFft& f, const float2* src, int64_t es, int64_t ls, float2* dst,
                                 int64_t tes, int64_t tls, int64_t count, const LargeEnds& e, float2* T,
                                 hipStream_t stream) const
{
    const int64_t M = f.m, M1 = int64_t(1) << f.logm1, M3 = int64_t(1) << f.logm3;
    const int64_t Mr = M / M1;
    enqueue_large_a1a2(f, src, es, ls, count, e, T, stream);
    // C: per (line, kr = k3 + M3*k2), element n1 at line*M + n1*M3 + k3 +
    // k2*M1*M3, times w_M^(n1*kr) on the load, the M1-point FFT over n1, stored
    // at kr + Mr*k1
    LargeJob c = large_job(f.tw_lo, f.tw_hi, e.inverse, e.n);
    c.src = T, c.dst = dst;
    c.sub_in = Mr, c.ls_in = M, c.sub2_in = M3, c.ls2_in = 1, c.ls3_in = M1 * M3, c.es_in = M3;
    c.sub_out = Mr, c.ls_out = tls, c.ls2_out = tes, c.es_out = Mr * tes;
    c.count = count * Mr;
    c.load_mode = LARGE_LOAD_TWIDDLE, c.tmod = Mr;
    c.store_mode = e.store_mode, c.kmul = Mr, c.scale = e.scale;
    c.chirp = e.chirp, c.dtab = e.dtab;
    large_step(f.logm1, twiddles(f.logm1), c, cu_count_, stream);
}

void Context::enqueue_large_blue3(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls,
                                  int64_t count, bool inverse, float2* T, hipStream_t stream) const
{
    const LargeFft f{lp.m, lp.logm1, lp.logm2, lp.logm3, lp.tw_lo, lp.tw_hi};
    const int64_t M = f.m, M1 = int64_t(1) << f.logm1, M2 = int64_t(1) << f.logm2, M3 = int64_t(1) << f.logm3;
    const int64_t Mr = M2 * M3;
    // the chirped line's Mr-point FFTs (x * c, zero past n)
    const LargeEnds e{LARGE_LOAD_CHIRP, LARGE_STORE_PLAIN, inverse, lp.n, 1.0f, lp.chirp, nullptr};
    enqueue_large_a1a2(f, src, es, ls, count, e, T, stream);
    // C fused, in place: per (line, kr) times w_M^(n1*kr), the M1-point FFT
    // over n1, conj(v * D) at kr + Mr*k1, the M1-point FFT over k1 (the second
    // transform with the factors' roles exchanged, j = j1 + M1*jr), times
    // w_M^(kr*j1), j1 stored in the slot of n1
    LargeJob c = large_job(f.tw_lo, f.tw_hi, inverse, lp.n);
    c.src = T, c.dst = T;
    c.sub_in = Mr, c.ls_in = M, c.sub2_in = M3, c.ls2_in = 1, c.ls3_in = M1 * M3, c.es_in = M3;
    c.sub_out = Mr, c.ls_out = M, c.sub2_out = M3, c.ls2_out = 1, c.ls3_out = M1 * M3, c.es_out = M3;
    c.count = count * Mr;
    c.load_mode = LARGE_LOAD_TWIDDLE, c.store_mode = LARGE_STORE_DIAG_FFT, c.tmod = Mr, c.kmul = Mr;
    c.dtab = lp.dtab;
    large_step(f.logm1, twiddles(f.logm1), c, cu_count_, stream);
    // X1, in place: per (line, j1*M3 + k3) the M2-point FFT over k2, times w_Mr^(k3*jr2) = w_M^(M1*k3*jr2)
    LargeJob x1 = large_job(f.tw_lo, f.tw_hi, inverse, lp.n);
    x1.src = T, x1.dst = T;
    x1.sub_in = M1 * M3, x1.ls_in = M, x1.ls2_in = 1, x1.es_in = M1 * M3;
    x1.sub_out = M1 * M3, x1.ls_out = M, x1.ls2_out = 1, x1.es_out = M1 * M3;
    x1.count = count * M1 * M3;
    x1.load_mode = LARGE_LOAD_PLAIN, x1.store_mode = LARGE_STORE_TWIDDLE;
    x1.tmod = M3, x1.tdiv = 1, x1.tmul = M1;
    large_step(f.logm2, twiddles(f.logm2), x1, cu_count_, stream);
    // X2: per (line, r = j1 + M1*jr2), element k3 at line*M + j1*M3 + jr2*M1*M3 + k3,
    // the M3-point FFT over k3 into j = r + M1*M2*jr3: conj(v) * c, crop to n, 1/n for the inverse
    LargeJob x2 = large_job(f.tw_lo, f.tw_hi, inverse, lp.n);
    x2.src = T, x2.dst = dst;
    x2.sub_in = M1 * M2, x2.ls_in = M, x2.sub2_in = M1, x2.ls2_in = M3, x2.ls3_in = M1 * M3, x2.es_in = 1;
    x2.sub_out = M1 * M2, x2.ls_out = ls, x2.ls2_out = es, x2.es_out = M1 * M2 * es;
    x2.count = count * M1 * M2;
    x2.load_mode = LARGE_LOAD_PLAIN, x2.store_mode = LARGE_STORE_CHIRP, x2.kmul = M1 * M2;
    x2.scale = inverse ? 1.0f / (float)lp.n : 1.0f;
    x2.chirp = lp.chirp;
    large_step(f.logm3, twiddles(f.logm3), x2, cu_count_, stream);
}

void Context::enqueue_large_fft(const LargeFft& lp, const float2* src, float2* dst, int64_t es, int64_t ls,
                                int64_t count, float scale, float2* scratch, hipStream_t stream) const
{
    const int64_t M = lp.m, M1 = int64_t(1) << lp.logm1, M2 = int64_t(1) << lp.logm2;
    if (lp.logm3 > 0) {
        // sign 0: the store of LARGE_STORE_CONJ is a plain store times scale
        const LargeEnds e{LARGE_LOAD_PLAIN, LARGE_STORE_CONJ, false, M, scale, nullptr, nullptr};
        enqueue_large_fft3(lp, src, es, ls, dst, es, ls, count, e, scratch, stream);
        return;
    }
    // forward FFT (power of two) of count lines: A then C
    LargeJob a = large_job(lp.tw_lo, lp.tw_hi, false, M);
    a.src = src;
    a.dst = scratch;
    a.sub_in = M1, a.ls_in = ls, a.ls2_in = es, a.es_in = M1 * es;
    a.sub_out = M1, a.ls_out = M, a.ls2_out = M2, a.es_out = 1;
    a.count = count * M1;
    a.load_mode = LARGE_LOAD_PLAIN, a.store_mode = LARGE_STORE_TWIDDLE;
    a.tmod = M1;  // w_M^(n1*k2)
    a.m1 = M1, a.m2 = M2;
    large_step(lp.logm2, twiddles(lp.logm2), a, cu_count_, stream);
    LargeJob c = large_job(lp.tw_lo, lp.tw_hi, false, M);
    c.src = scratch;
    c.dst = dst;
    c.sub_in = M2, c.ls_in = M, c.ls2_in = 1, c.es_in = M2;
    c.sub_out = M2, c.ls_out = ls, c.ls2_out = es, c.es_out = M2 * es;
    c.count = count * M2;
    c.load_mode = LARGE_LOAD_PLAIN, c.store_mode = LARGE_STORE_CONJ;  // sign 0: a plain store times scale
    c.m1 = M1, c.m2 = M2, c.scale = scale;
    large_step(lp.logm1, twiddles(lp.logm1), c, cu_count_, stream);
}

void Context::enqueue_large_axis(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls,
                                 int64_t count, bool inverse, float2* scratch, hipStream_t stream) const
{
    const int64_t M = lp.m, M1 = int64_t(1) << lp.logm1, M2 = int64_t(1) << lp.logm2;
    const float scale = inverse ? 1.0f / (float)lp.n : 1.0f;
    if (lp.logm3 > 0) {
        const LargeFft f{M, lp.logm1, lp.logm2, lp.logm3, lp.tw_lo, lp.tw_hi};
        if (!lp.blue) {
            // conj(x) for the inverse on the load, conj(v)/n on the store
            const LargeEnds e{LARGE_LOAD_CONJ, LARGE_STORE_CONJ, inverse, lp.n, scale, nullptr, nullptr};
            enqueue_large_fft3(f, src, es, ls, dst, es, ls, count, e, scratch, stream);
            return;
        }
        enqueue_large_blue3(lp, src, dst, es, ls, count, inverse, scratch, stream);
        return;
    }
    float2* S = scratch;  // step A output, count * M (later steps in place or out of it)
    auto step_a = [&](const float2* from, int64_t fls, int64_t fes, int load_mode) {
        LargeJob a = large_job(lp.tw_lo, lp.tw_hi, inverse, lp.n);
        a.src = from;
        a.dst = S;
        a.sub_in = M1, a.ls_in = fls, a.ls2_in = fes, a.es_in = M1 * fes;
        a.sub_out = M1, a.ls_out = M, a.ls2_out = M2, a.es_out = 1;
        a.count = count * M1;
        a.load_mode = load_mode, a.store_mode = LARGE_STORE_TWIDDLE;
        a.tmod = M1;  // w_M^(n1*k2)
        a.m1 = M1, a.m2 = M2;
        a.chirp = lp.chirp;
        large_step(lp.logm2, twiddles(lp.logm2), a, cu_count_, stream);
    };
    auto step_c = [&](float2* to, int64_t tls, int64_t tes, int store_mode, float sc) {
        LargeJob c = large_job(lp.tw_lo, lp.tw_hi, inverse, lp.n);
        c.src = S;
        c.dst = to;
        c.sub_in = M2, c.ls_in = M, c.ls2_in = 1, c.es_in = M2;
        c.sub_out = M2, c.ls_out = tls, c.ls2_out = tes, c.es_out = M2 * tes;
        c.count = count * M2;
        c.load_mode = LARGE_LOAD_PLAIN, c.store_mode = store_mode;
        c.m1 = M1, c.m2 = M2, c.kmul = M2, c.scale = sc;
        c.chirp = lp.chirp, c.dtab = lp.dtab;
        large_step(lp.logm1, twiddles(lp.logm1), c, cu_count_, stream);
    };
    if (!lp.blue) {
        step_a(src, ls, es, LARGE_LOAD_CONJ);         // conj(x) for the inverse
        step_c(dst, ls, es, LARGE_STORE_CONJ, scale);  // conj(v)/n for the inverse
        return;
    }
    step_a(src, ls, es, LARGE_LOAD_CHIRP);  // x * c, zero past n
    // B, in place: per (line, k2) the M1-point FFT over n1, conj(v * D), the
    // M1-point FFT over k1 and the twiddle w_M^(k2*j1), at j1*M2 + k2
    LargeJob b = large_job(lp.tw_lo, lp.tw_hi, inverse, lp.n);
    b.src = S;
    b.dst = S;
    b.sub_in = M2, b.ls_in = M, b.ls2_in = 1, b.es_in = M2;
    b.sub_out = M2, b.ls_out = M, b.ls2_out = 1, b.es_out = M2;
    b.count = count * M2;
    b.load_mode = LARGE_LOAD_PLAIN, b.store_mode = LARGE_STORE_DIAG_FFT;
    b.tmod = M2;  // w_M^(k2*j1)
    b.m1 = M1, b.m2 = M2, b.kmul = M2;
    b.dtab = lp.dtab;
    large_step(lp.logm1, twiddles(lp.logm1), b, cu_count_, stream);
    // C', per (line, j1): the M2-point FFT over k2 into j = j1 + M1*j2:
    // conj(v) * c, crop to n, 1/n for the inverse
    LargeJob c = large_job(lp.tw_lo, lp.tw_hi, inverse, lp.n);
    c.src = S;
    c.dst = dst;
    c.sub_in = M1, c.ls_in = M, c.ls2_in = M2, c.es_in = 1;
    c.sub_out = M1, c.ls_out = ls, c.ls2_out = es, c.es_out = M1 * es;
    c.count = count * M1;
    c.load_mode = LARGE_LOAD_PLAIN, c.store_mode = LARGE_STORE_CHIRP;
    c.m1 = M1, c.m2 = M2, c.kmul = M1, c.scale = scale;
    c.chirp = lp.chirp;
    large_step(lp.logm2, twiddles(lp.logm2), c, cu_count_, stream);
}

void Context::enqueue_axis(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls, int count,
                           bool inverse, bool cols, float2* scratch, hipStream_t stream) const
{
    if (lp.large()) {
        if (!scratch) throw Error(MDFT_ERR_STATE, "long-line pass without a scratch buffer");
        enqueue_large_axis(lp, src, dst, es, ls, count, inverse, scratch, stream);
        return;
    }
    const AxisJob j = make_job(lp, src, dst, es, ls, count, inverse);
    hip_check(launch_axis(lp.shape, j, lp.blue, cols, tile_threads(lp.shape, lp.blue, cols, count), stream),
              "axis pass launch");
}

bool Context::wide_tiles(int shape, bool blue, bool cols, int count) const
{
    if (variant(shape, blue, cols) == 0) return false;
    // wide tiles hold 4x the lines of narrow ones; keep at least one tile per CU
    const int tpl = m_of(shape) / ept_of(shape);
    const int lines_wide = tpl >= 1024 ? 1 : 1024 / tpl;
    const int64_t tiles = (int64_t(count) + lines_wide - 1) / lines_wide;
    return tiles >= int64_t(cu_count_);
}

int Context::tile_threads(int shape, bool blue, bool cols, int count) const
{
    if (wide_tiles(shape, blue, cols, count)) return 1024;
    // passes with few lines: smaller tiles (one wave for lines of up to 1024
    // points) so every CU gets work; rows aim at four tiles per CU, columns at
    // one, since narrower column tiles read shorter runs of each row
    const int tpl = m_of(shape) / ept_of(shape);
    const int64_t want = int64_t(cols ? 1 : 4) * cu_count_;
    int t = tpl > 256 ? tpl : 256;
    while (t > 64 && t / 2 >= tpl && (int64_t(count) + t / tpl - 1) / (t / tpl) < want) t /= 2;
    return t;
}

void Context::enqueue_transform(const LinePlan& row, const LinePlan& col, int w, int h, bool inverse,
                                const float2* src, float2* dst, float2* scratch, hipStream_t stream) const
{
    const float2* s = src;
    if (w > 1) {
        enqueue_axis(row, s, dst, 1, w, h, inverse, false, scratch, stream);
        s = dst;
    }
    if (h > 1) {
        enqueue_axis(col, s, dst, w, 1, w, inverse, true, scratch, stream);
        s = dst;
    }
    if (s != dst)
        hip_check(hipMemcpyAsync(dst, src, sizeof(float2) * (size_t)w * h, hipMemcpyDeviceToDevice, stream),
                  "identity copy");
}

}  // namespace mdft
