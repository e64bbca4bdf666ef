// C ABI of the single-transform path: plans, host and device exec, errors.
#include "capi_util.hpp"
#include "staging.hpp"

#include <algorithm>
#include <cstring>
#include <string>

using mdft::Context;
using mdft::Error;
using mdft::LinePlan;

struct mdft_plan_s {
    int device = 0;
    uint32_t w = 0, h = 0;
    std::shared_ptr<const LinePlan> row, col;
    float2* xd = nullptr;          // device image for mdft_exec, allocated on first use
    float2* scratch = nullptr;     // long-line passes only
    uint64_t scratch_elems = 0;
    hipStream_t stream = nullptr;  // private stream of mdft_exec
    mdft::HostStaging staging;     // pinned ring of mdft_exec
    mdft::ScratchOrder order;      // calls sharing `scratch` on different streams
};

namespace mdft {
namespace {
thread_local std::string t_error;
}
void set_error(const std::string& s) { t_error = s; }
const char* get_error() { return t_error.c_str(); }
}  // namespace mdft

extern "C" {

const char* mdft_last_error(void) { return mdft::get_error(); }

const char* mdft_version(void) { return "mdft 0.1 (gfx950)"; }

int mdft_init(int device)
{
    MDFT_API_BEGIN
    Context::get(device);
    Context::set_thread_device(device);
    MDFT_API_END
}

int mdft_plan_create(mdft_plan* plan, uint32_t width, uint32_t height)
{
    MDFT_API_BEGIN
    if (!plan) throw Error(MDFT_ERR_INVALID, "mdft_plan_create: plan is NULL");
    *plan = nullptr;
    // validate both sides with the reference's message before touching the device
    mdft::compute_m((long long)width);
    mdft::compute_m((long long)height);
    Context& ctx = Context::get(Context::current_device());
    auto p = std::make_unique<mdft_plan_s>();
    p->device = ctx.device();
    p->w = width;
    p->h = height;
    p->row = ctx.line_plan((int)width, false);
    p->col = ctx.line_plan((int)height, true);
    p->scratch_elems = std::max(width > 1 ? p->row->scratch_elems((int)height) : uint64_t(0),
                                height > 1 ? p->col->scratch_elems((int)width) : uint64_t(0));
    if (p->scratch_elems) {
        mdft::DeviceGuard g(p->device);
        mdft::hip_check(hipMalloc(&p->scratch, sizeof(float2) * p->scratch_elems), "hipMalloc scratch");
    }
    *plan = p.release();
    MDFT_API_END
}

int mdft_plan_destroy(mdft_plan plan)
{
    MDFT_API_BEGIN
    if (!plan) return MDFT_OK;
    {
        mdft::DeviceGuard g(plan->device);
        if (plan->stream) {
            (void)hipStreamSynchronize(plan->stream);
            (void)hipStreamDestroy(plan->stream);
        }
        if (plan->xd) (void)hipFree(plan->xd);
        if (plan->scratch) (void)hipFree(plan->scratch);
        plan->order.release();
        plan->staging.release();
    }
    delete plan;
    MDFT_API_END
}

int mdft_plan_info(mdft_plan plan, uint32_t* m1, uint32_t* m2, uint64_t* device_bytes)
{
    MDFT_API_BEGIN
    if (!plan) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (m1) *m1 = (uint32_t)plan->row->padded;
    if (m2) *m2 = (uint32_t)plan->col->padded;
    if (device_bytes)
        *device_bytes = plan->row->bytes + (plan->col != plan->row ? plan->col->bytes : 0) +
                        (plan->xd ? sizeof(float2) * (uint64_t)plan->w * plan->h : 0) +
                        sizeof(float2) * plan->scratch_elems;
    MDFT_API_END
}

int mdft_exec(mdft_plan plan, int inverse, void* data, size_t count)
{
    MDFT_API_BEGIN
    // src/gpu/dft/compute.cpp:259-267
    if (!plan || plan->w == 0 || plan->h == 0) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!data || count != (size_t)plan->w * plan->h) throw Error(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size");
    mdft::DeviceGuard g(plan->device);
    const size_t bytes = sizeof(float2) * count;
    if (!plan->stream)
        mdft::hip_check(hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!plan->xd) mdft::hip_check(hipMalloc(&plan->xd, bytes), "hipMalloc image");
    Context& ctx = Context::get(plan->device);
    plan->staging.upload(plan->xd, data, bytes, plan->stream);
    if (plan->scratch) plan->order.before(plan->stream);
    ctx.enqueue_transform(*plan->row, *plan->col, (int)plan->w, (int)plan->h, inverse != 0, plan->xd, plan->xd,
                          plan->scratch, plan->stream);
    if (plan->scratch) plan->order.after(plan->stream);
    plan->staging.download(data, plan->xd, bytes, plan->stream);
    mdft::hip_check(hipStreamSynchronize(plan->stream), "exec");
    MDFT_API_END
}

int mdft_exec_device(mdft_plan plan, int inverse, const void* src, void* dst, void* stream)
{
    MDFT_API_BEGIN
    if (!plan || plan->w == 0 || plan->h == 0) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!src || !dst) throw Error(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size");
    mdft::DeviceGuard g(plan->device);
    Context& ctx = Context::get(plan->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (plan->scratch) plan->order.before(s);
    ctx.enqueue_transform(*plan->row, *plan->col, (int)plan->w, (int)plan->h, inverse != 0,
                          static_cast<const float2*>(src), static_cast<float2*>(dst), plan->scratch, s);
    if (plan->scratch) plan->order.after(s);
    MDFT_API_END
}

int mdft_cache_info(uint64_t* bytes, uint32_t* lengths, uint64_t* limit)
{
    MDFT_API_BEGIN
    Context& ctx = Context::get(Context::current_device());
    if (bytes) *bytes = ctx.cached_bytes();
    if (lengths) *lengths = ctx.cached_lengths();
    if (limit) *limit = ctx.cache_limit();
    MDFT_API_END
}

int mdft_cache_set_limit(uint64_t bytes)
{
    MDFT_API_BEGIN
    Context::get(Context::current_device()).set_cache_limit(bytes);
    MDFT_API_END
}

static void check_tune_args(int shape, int blue, int cols)
{
    const bool ok_shape = (shape >= 1 && shape <= mdft::MAX_LOGM) ||
                          (mdft::is_f3(shape) && mdft::f3_k(shape) >= mdft::F3_MIN_K &&
                           mdft::f3_k(shape) <= mdft::F3_MAX_K);
    if (!ok_shape || (blue != 0 && blue != 1) || (cols != 0 && cols != 1))
        throw Error(MDFT_ERR_INVALID, "tune: shape must be in [1, 14] or [38, 43], blue and cols 0 or 1");
}

int mdft_tune_get(int logm, int blue, int cols, int* variant)
{
    MDFT_API_BEGIN
    check_tune_args(logm, blue, cols);
    if (!variant) throw Error(MDFT_ERR_INVALID, "tune: variant is NULL");
    *variant = Context::get(Context::current_device()).variant(logm, blue != 0, cols != 0);
    MDFT_API_END
}

int mdft_tune_set(int logm, int blue, int cols, int variant)
{
    MDFT_API_BEGIN
    check_tune_args(logm, blue, cols);
    if (variant != 0 && variant != 1) throw Error(MDFT_ERR_INVALID, "tune: variant must be 0 or 1");
    Context::get(Context::current_device()).set_variant(logm, blue != 0, cols != 0, variant);
    MDFT_API_END
}

}  // extern "C"
