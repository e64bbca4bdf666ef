// C ABI of the batched power-of-two 1-D FFT (count lines of length n, back
// to back), the reference's Fft component (src/gpu/dft/fft.h:32-53,
// fft.cpp:207-330) on this library's row kernels.
#include "capi_util.hpp"
#include "staging.hpp"

#include <string>

using mdft::Context;
using mdft::Error;
using mdft::LinePlan;

struct mdft_fft_s {
    int device = 0;
    uint32_t count = 0, n = 0;
    std::shared_ptr<const LinePlan> line;  // null for n == 1
    float2* xd = nullptr;                  // device copy for mdft_fft_exec, allocated on first use
    float2* scratch = nullptr;             // lines longer than one tile
    hipStream_t stream = nullptr;
    mdft::HostStaging staging;             // pinned ring of mdft_fft_exec
    mdft::ScratchOrder order;              // calls sharing `scratch` on different streams
};

extern "C" {

int mdft_fft_create(mdft_fft* fft, uint32_t count, uint32_t n)
{
    MDFT_API_BEGIN
    if (!fft) throw Error(MDFT_ERR_INVALID, "mdft_fft_create: fft is NULL");
    *fft = nullptr;
    // src/gpu/dft/fft.cpp:298-306
    if (n == 0) throw Error(MDFT_ERR_SIZE, "FFT size " + std::to_string(n) + " is not positive");
    if (!mdft::is_pow2(n))
        throw Error(MDFT_ERR_SIZE, "FFT size " + std::to_string(n) + " is not an integral power of 2");
    // the longest line of a 2-D plan (compute_m's limit, function.cpp:34-38)
    if (n > (1u << 29))
        throw Error(MDFT_ERR_SIZE, "FFT size " + std::to_string(n) + " exceeds the longest supported line (2^29)");
    if (count > 0x7fffffffu || (uint64_t)count * n > (uint64_t(1) << 40))
        throw Error(MDFT_ERR_SIZE, "FFT count " + std::to_string(count) + " x size " + std::to_string(n) +
                                       " is too large");
    Context& ctx = Context::get(Context::current_device());
    auto f = std::make_unique<mdft_fft_s>();
    f->device = ctx.device();
    f->count = count;
    f->n = n;
    if (n > 1) {
        f->line = ctx.line_plan((int)n, false);
        const uint64_t se = f->line->scratch_elems((int)count);
        if (se) {
            mdft::DeviceGuard g(f->device);
            mdft::hip_check(hipMalloc(&f->scratch, sizeof(float2) * se), "hipMalloc scratch");
        }
    }
    *fft = f.release();
    MDFT_API_END
}

int mdft_fft_destroy(mdft_fft fft)
{
    MDFT_API_BEGIN
    if (!fft) return MDFT_OK;
    {
        mdft::DeviceGuard g(fft->device);
        if (fft->stream) {
            (void)hipStreamSynchronize(fft->stream);
            (void)hipStreamDestroy(fft->stream);
        }
        if (fft->xd) (void)hipFree(fft->xd);
        if (fft->scratch) (void)hipFree(fft->scratch);
        fft->order.release();
        fft->staging.release();
    }
    delete fft;
    MDFT_API_END
}

int mdft_fft_exec(mdft_fft fft, int inverse, void* data, size_t size)
{
    MDFT_API_BEGIN
    if (!fft) throw Error(MDFT_ERR_NO_BUFFERS, "No FFT plan");
    // n == 1: the reference has no shader objects and returns before its size check (fft.cpp:271-274)
    if (fft->n == 1) return MDFT_OK;
    const uint64_t need = (uint64_t)fft->count * fft->n;
    if (!data || size != need)
        throw Error(MDFT_ERR_DATA_SIZE, "FFT buffer size " + std::to_string(size) + " must be equal to data size " +
                                            std::to_string(need));
    if (need == 0) return MDFT_OK;
    mdft::DeviceGuard g(fft->device);
    const size_t bytes = sizeof(float2) * size;
    if (!fft->stream) mdft::hip_check(hipStreamCreateWithFlags(&fft->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!fft->xd) mdft::hip_check(hipMalloc(&fft->xd, bytes), "hipMalloc data");
    Context& ctx = Context::get(fft->device);
    fft->staging.upload(fft->xd, data, bytes, fft->stream);
    if (fft->scratch) fft->order.before(fft->stream);
    ctx.enqueue_axis(*fft->line, fft->xd, fft->xd, 1, fft->n, (int)fft->count, inverse != 0, false, fft->scratch,
                     fft->stream);
    if (fft->scratch) fft->order.after(fft->stream);
    fft->staging.download(data, fft->xd, bytes, fft->stream);
    mdft::hip_check(hipStreamSynchronize(fft->stream), "exec");
    MDFT_API_END
}

int mdft_fft_exec_device(mdft_fft fft, int inverse, const void* src, void* dst, void* stream)
{
    MDFT_API_BEGIN
    if (!fft) throw Error(MDFT_ERR_NO_BUFFERS, "No FFT plan");
    if (!src || !dst) throw Error(MDFT_ERR_DATA_SIZE, "FFT buffer is NULL");
    const uint64_t need = (uint64_t)fft->count * fft->n;
    if (need == 0) return MDFT_OK;
    mdft::DeviceGuard g(fft->device);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (fft->n == 1) {
        if (src != dst)
            mdft::hip_check(hipMemcpyAsync(dst, src, sizeof(float2) * need, hipMemcpyDeviceToDevice, s), "copy");
        return MDFT_OK;
    }
    if (fft->scratch) fft->order.before(s);
    Context::get(fft->device)
        .enqueue_axis(*fft->line, static_cast<const float2*>(src), static_cast<float2*>(dst), 1, fft->n,
                      (int)fft->count, inverse != 0, false, fft->scratch, s);
    if (fft->scratch) fft->order.after(s);
    MDFT_API_END
}

}  // extern "C"
