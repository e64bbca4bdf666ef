// Per-device context: twiddle tables, the cache of per-length line plans
// (chirp and diagonal tables generated in FP64 on the device), and the
// enqueueing of axis passes.
#include "runtime.hpp"

#include "../../../include/mdft.h"

#include <string>
#include <vector>

namespace mdft {

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

LinePlan::~LinePlan()
{
    if (chirp) (void)hipFree(chirp);
    if (dtab) (void)hipFree(dtab);
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

namespace {
thread_local int t_device = -1;
std::mutex g_ctx_mu;
std::unordered_map<int, Context*> g_ctx;
}  // namespace

int Context::current_device() { return t_device >= 0 ? t_device : 0; }

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
    // default tile shapes: wide tiles for columns of 1024 points and more
    for (int l = 0; l <= MAX_LOGM; ++l)
        for (int b = 0; b < 2; ++b) variant_[l][b][1] = (l >= 10) ? 1 : 0;
}

std::shared_ptr<const LinePlan> Context::line_plan(int n)
{
    const int m = compute_m(n);  // validates n
    std::lock_guard<std::mutex> lk(mu_);
    auto it = lines_.find(n);
    if (it != lines_.end()) return it->second;

    auto lp = std::make_shared<LinePlan>();
    lp->n = n;
    lp->m = m;
    lp->logm = ilog2((uint64_t)m);
    lp->blue = !is_pow2((uint64_t)n);
    if (lp->logm > MAX_LOGM)
        throw Error(MDFT_ERR_SIZE, "Error size " + std::to_string(n) + ": padded length " + std::to_string(m) +
                                       " exceeds the longest supported line (" + std::to_string(1 << MAX_LOGM) + ")");
    if (lp->blue) {
        DeviceGuard g(device_);
        const size_t need = 2 * (size_t)n + (size_t)m;
        if (need > scratch64_doubles_) {
            if (scratch64_) hip_check(hipFree(scratch64_), "hipFree");
            scratch64_ = nullptr;
            hip_check(hipMalloc(&scratch64_, sizeof(double) * need), "hipMalloc scratch");
            scratch64_doubles_ = need;
        }
        hip_check(hipMalloc(&lp->chirp, sizeof(float2) * (size_t)n), "hipMalloc chirp");
        hip_check(hipMalloc(&lp->dtab, sizeof(float2) * (size_t)m), "hipMalloc dtab");
        lp->bytes = sizeof(float2) * ((size_t)n + (size_t)m);
        double2* chirp64 = reinterpret_cast<double2*>(scratch64_);
        hip_check(gen_chirp(chirp64, lp->chirp, n, gen_stream_), "chirp generation");
        hip_check(gen_dtab(lp->dtab, chirp64, scratch64_ + 2 * (size_t)n, n, m, gen_stream_), "diagonal generation");
        hip_check(hipStreamSynchronize(gen_stream_), "table generation");
    }
    lines_[n] = lp;
    return lp;
}

uint64_t Context::cached_bytes()
{
    std::lock_guard<std::mutex> lk(mu_);
    uint64_t s = sizeof(float2) * (uint64_t)tw_total();
    for (auto& kv : lines_) s += kv.second->bytes;
    return s;
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
    j.chirp = lp.chirp;
    j.dtab = lp.dtab;
    j.tw = twiddles(lp.logm);
    return j;
}

void Context::enqueue_axis(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls, int count,
                           bool inverse, bool cols, hipStream_t stream) const
{
    const AxisJob j = make_job(lp, src, dst, es, ls, count, inverse);
    hip_check(launch_axis(lp.logm, j, lp.blue, cols, variant(lp.logm, lp.blue, cols) != 0, stream),
              "axis pass launch");
}

void Context::enqueue_transform(const LinePlan& row, const LinePlan& col, int w, int h, bool inverse,
                                const float2* src, float2* dst, hipStream_t stream) const
{
    const float2* s = src;
    if (w > 1) {
        enqueue_axis(row, s, dst, 1, w, h, inverse, false, stream);
        s = dst;
    }
    if (h > 1) {
        enqueue_axis(col, s, dst, w, 1, w, inverse, true, stream);
        s = dst;
    }
    if (s != dst)
        hip_check(hipMemcpyAsync(dst, src, sizeof(float2) * (size_t)w * h, hipMemcpyDeviceToDevice, stream),
                  "identity copy");
}

}  // namespace mdft
