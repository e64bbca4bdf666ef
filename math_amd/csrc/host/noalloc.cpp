// Allocation-free transform entry point (mdft_exec_device_noalloc) for
// callers that cannot allocate device memory: code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc.  It runs the same
// kernels as mdft_exec_device; the scratch buffer that lines longer than one
// tile need lives in a static device array instead of plan-owned memory.
//
// The array's name is a host-side stand-in: kernels get its device address,
// looked up once per device with hipGetSymbolAddress when the device's
// context is created (mdft_init or the first plan), and never the name itself.
// Every call on a device shares that device's array and leaves its contents
// behind: calls that overlap on one GPU (two streams, or a replayed graph
// beside an eager call) race on it.  Nothing in it has to start at zero: the
// four-step passes write every scratch element they later read.
#include "capi_util.hpp"

#include <algorithm>
#include <mutex>
#include <string>

using mdft::Context;
using mdft::Error;

namespace {

// 4 Mi complex<float> = 32 MiB per device, held until the process ends
constexpr uint64_t kScratchElems = uint64_t(4) << 20;
static __device__ __align__(256) float2 scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float2* g_ptr[kMaxDevices] = {};

}  // namespace

namespace mdft {

void noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) throw Error(MDFT_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return;
    // the caller (context creation) has made `device` current
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch");
    g_ptr[device] = static_cast<float2*>(p);
}

float2* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

uint64_t noalloc_scratch_elems() { return kScratchElems; }

}  // namespace mdft

namespace {

// Scratch elements one axis pass of `count` lines of length n needs: lines
// whose padded length exceeds one tile run the four-step path through
// count * M elements.
uint64_t pass_scratch(uint32_t n, uint32_t count)
{
    const int m = mdft::compute_m(n);  // "Error size <n> for compute m" outside [1, 2^29]
    if (n <= 1) return 0;
    return mdft::large_scratch_elems(m, count);
}

}  // namespace

extern "C" {

int mdft_exec_device_noalloc(uint32_t width, uint32_t height, int inverse, const void* src, void* dst, void* stream)
{
    MDFT_API_BEGIN
    if (!src || !dst) throw Error(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size");
    const uint64_t need = std::max(pass_scratch(width, height), pass_scratch(height, width));
    if (need > kScratchElems)
        throw Error(MDFT_ERR_SIZE, "mdft_exec_device_noalloc: a " + std::to_string(width) + " x " +
                                       std::to_string(height) + " transform needs " + std::to_string(need) +
                                       " scratch elements, above the static limit of " +
                                       std::to_string(kScratchElems));
    int device = -1;
    mdft::hip_check(hipGetDevice(&device), "hipGetDevice");
    Context* ctx = Context::find(device);
    float2* ptr = mdft::noalloc_scratch(device);
    if (!ctx || !ptr)
        throw Error(MDFT_ERR_STATE, "mdft_exec_device_noalloc: device " + std::to_string(device) +
                                        " is not initialised (call mdft_init or create a plan first)");
    auto row = ctx->find_line_plan((int)width, false);
    auto col = ctx->find_line_plan((int)height, true);
    if (!row || !col)
        throw Error(MDFT_ERR_STATE, "mdft_exec_device_noalloc: the tables of length " +
                                        std::to_string(!row ? width : height) +
                                        " are not resident (keep a plan of this size alive)");
    ctx->enqueue_transform(*row, *col, (int)width, (int)height, inverse != 0, static_cast<const float2*>(src),
                           static_cast<float2*>(dst), ptr, static_cast<hipStream_t>(stream));
    MDFT_API_END
}

}  // extern "C"
