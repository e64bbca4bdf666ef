// C ABI of batched transforms of mixed sizes (job tables and work lists are
// built by batch_builder.cpp).
#include "batch_builder.hpp"
#include "capi_util.hpp"

#include <algorithm>
#include <numeric>
#include <tuple>
#include <vector>

using mdft::BatchJob;
using mdft::Context;
using mdft::Error;
using mdft::LinePlan;

struct mdft_batch_s {
    int device = 0;
    uint32_t flags = 0;
    int threads = 256;       // whole images and row tiles
    int threads_cols = 256;  // column tiles
    std::vector<mdft_desc> descs;
    mdft::BatchBuild build;
    std::vector<uint32_t> job_image;
    std::vector<uint8_t> job_is_col;
    std::vector<uint32_t> copies;  // 1x1 identities (image index)
    uint32_t nphase1 = 0, nitems = 0;
    BatchJob* jobs_dev = nullptr;
    BatchJob* jobs_pinned = nullptr;
    uint2* items_dev = nullptr;
    uint32_t* ctl_dev = nullptr;  // two control blocks of nctl words
    uint32_t nctl = 0;
    int grid = 0;  // persistent grid
    const void* cur_src = nullptr;
    void* cur_dst = nullptr;
    hipStream_t last_stream = nullptr;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t spin_limit = 1u << 20;  // fail-safe of the dependency wait: then the error flag
    uint64_t phase_bytes[2] = {0, 0};  // algorithmic bytes (16 per point and pass) per launch phase
};

namespace {

// Algorithmic bytes of each launch phase: every axis pass reads and writes
// its image once (16 bytes per point); a job belongs to the phase of its items.
void count_phase_bytes(mdft_batch_s* b)
{
    const auto& items = b->build.items;
    std::vector<int> phase(b->build.jobs.size(), 0);
    for (size_t i = 0; i < items.size(); ++i) {
        const int ph = i < b->nphase1 ? 1 : 2;
        phase[items[i].x] = ph;
        if (items[i].y & mdft::BATCH_WHOLE) phase[items[i].y & ~mdft::BATCH_WHOLE] = ph;
    }
    for (size_t j = 0; j < phase.size(); ++j) {
        const mdft_desc& d = b->descs[b->job_image[j]];
        if (phase[j]) b->phase_bytes[phase[j] - 1] += 16ull * d.width * d.height;
    }
}

void enqueue_batch(mdft_batch_s* b, hipStream_t stream)
{
    const int T = b->threads;
    if (b->flags & MDFT_BATCH_DEPENDENT) {
        mdft::hip_check(mdft::launch_batch_persistent(T, b->jobs_dev, b->items_dev, b->nitems, b->ctl_dev, b->nctl,
                                                      b->grid, b->spin_limit, stream),
                        "batch launch");
    } else if (b->flags & MDFT_BATCH_PERSISTENT) {
        mdft::hip_check(mdft::launch_batch_persistent(T, b->jobs_dev, b->items_dev, b->nphase1, b->ctl_dev, b->nctl,
                                                      b->grid, b->spin_limit, stream),
                        "batch launch (phase 1)");
        mdft::hip_check(mdft::launch_batch_persistent(T, b->jobs_dev, b->items_dev + b->nphase1,
                                                      b->nitems - b->nphase1, b->ctl_dev + b->nctl, b->nctl, b->grid,
                                                      b->spin_limit, stream),
                        "batch launch (phase 2)");
    } else {
        mdft::hip_check(mdft::launch_batch_grid(T, b->jobs_dev, b->items_dev, b->nphase1, stream),
                        "batch launch (phase 1)");
        mdft::hip_check(mdft::launch_batch_grid(b->threads_cols, b->jobs_dev, b->items_dev + b->nphase1,
                                                b->nitems - b->nphase1, stream),
                        "batch launch (phase 2)");
    }
}

// Time the column phase of one batch layout (column tiles of TC threads) on
// zeroed scratch images: build, upload, one warm-up and three timed launches.
float time_column_phase(Context& ctx, const std::vector<mdft::BatchEntry>& entries, int T, int TC, hipStream_t s)
{
    mdft::BatchBuild bb;
    mdft::build_batch(ctx, entries.data(), entries.size(), T, TC, false, bb);
    const uint32_t n2 = (uint32_t)bb.items.size() - bb.nphase1;
    if (n2 == 0) return 0.0f;
    BatchJob* jobs = nullptr;
    uint2* items = nullptr;
    hipEvent_t a = nullptr, e = nullptr;
    float ms = 0.0f;
    hipError_t err = hipMalloc(&jobs, sizeof(BatchJob) * bb.jobs.size());
    if (err == hipSuccess) err = hipMalloc(&items, sizeof(uint2) * n2);
    if (err == hipSuccess)
        err = hipMemcpy(jobs, bb.jobs.data(), sizeof(BatchJob) * bb.jobs.size(), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(items, bb.items.data() + bb.nphase1, sizeof(uint2) * n2, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipEventCreate(&a);
    if (err == hipSuccess) err = hipEventCreate(&e);
    if (err == hipSuccess) err = mdft::launch_batch_grid(TC, jobs, items, n2, s);
    if (err == hipSuccess) err = hipEventRecord(a, s);
    for (int r = 0; r < 3 && err == hipSuccess; ++r) err = mdft::launch_batch_grid(TC, jobs, items, n2, s);
    if (err == hipSuccess) err = hipEventRecord(e, s);
    if (err == hipSuccess) err = hipEventSynchronize(e);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, e);
    (void)hipStreamSynchronize(s);
    if (a) (void)hipEventDestroy(a);
    if (e) (void)hipEventDestroy(e);
    if (jobs) (void)hipFree(jobs);
    if (items) (void)hipFree(items);
    mdft::hip_check(err, "batch tuning");
    return ms / 3.0f;
}

// Column-tile size of a grid-shape batch picked by measurement: 256 (when the
// row tiles are 256-thread), 512 or 1024 threads.
int tune_column_threads(Context& ctx, const mdft_desc* descs, uint32_t count, int T)
{
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; ++i) total += (uint64_t)descs[i].width * descs[i].height;
    float2* scratch = nullptr;
    mdft::hip_check(hipMalloc(&scratch, sizeof(float2) * std::max<uint64_t>(total, 1)), "hipMalloc tuning images");
    hipStream_t s = nullptr;
    int best = mdft::BATCH_COL_THREADS > T ? mdft::BATCH_COL_THREADS : T;
    try {
        mdft::hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        mdft::hip_check(hipMemsetAsync(scratch, 0, sizeof(float2) * std::max<uint64_t>(total, 1), s), "clear");
        std::vector<mdft::BatchEntry> entries(count);
        uint64_t off = 0;
        for (uint32_t i = 0; i < count; ++i) {
            entries[i] = {descs[i].width, descs[i].height, descs[i].inverse != 0, scratch + off, scratch + off};
            off += (uint64_t)descs[i].width * descs[i].height;
        }
        float best_ms = -1.0f;
        for (int tc : {256, 512, 1024}) {
            if (tc < T) continue;
            const float ms = time_column_phase(ctx, entries, T, tc, s);
            if (best_ms < 0 || ms < best_ms) {
                best_ms = ms;
                best = tc;
            }
        }
    } catch (...) {
        if (s) (void)hipStreamDestroy(s);
        (void)hipFree(scratch);
        throw;
    }
    (void)hipStreamDestroy(s);
    (void)hipFree(scratch);
    return best;
}

}  // namespace

extern "C" {

int mdft_batch_create(mdft_batch* batch, const mdft_desc* descs, uint32_t count, uint32_t flags)
{
    MDFT_API_BEGIN
    if (!batch) throw Error(MDFT_ERR_INVALID, "mdft_batch_create: batch is NULL");
    *batch = nullptr;
    if (!descs && count) throw Error(MDFT_ERR_INVALID, "mdft_batch_create: descs is NULL");
    std::vector<uint32_t> ws(count), hs(count);
    for (uint32_t i = 0; i < count; ++i) {
        ws[i] = descs[i].width;
        hs[i] = descs[i].height;
    }
    const int T = mdft::batch_threads_for(ws.data(), hs.data(), count);  // validates sizes

    Context& ctx = Context::get(Context::current_device());
    mdft::DeviceGuard g(ctx.device());
    auto b = std::make_unique<mdft_batch_s>();
    b->device = ctx.device();
    b->flags = flags;
    b->threads = T;
    // grid launches run column tiles in wide tiles; the persistent shapes use one size
    const bool grid_shape = !(flags & (MDFT_BATCH_PERSISTENT | MDFT_BATCH_DEPENDENT));
    b->threads_cols = grid_shape ? (mdft::BATCH_COL_THREADS > T ? mdft::BATCH_COL_THREADS : T) : T;
    if (grid_shape && (flags & MDFT_BATCH_TUNE)) b->threads_cols = tune_column_threads(ctx, descs, count, T);
    b->descs.assign(descs, descs + count);
    std::vector<mdft::BatchEntry> entries(count);
    for (uint32_t i = 0; i < count; ++i) {
        entries[i] = {descs[i].width, descs[i].height, descs[i].inverse != 0, nullptr, nullptr};
        if (descs[i].width == 1 && descs[i].height == 1) b->copies.push_back(i);
    }
    mdft::build_batch(ctx, entries.data(), count, T, b->threads_cols, (flags & MDFT_BATCH_DEPENDENT) != 0, b->build,
                      &b->job_image, &b->job_is_col);
    b->nphase1 = b->build.nphase1;
    b->nitems = (uint32_t)b->build.items.size();
    b->nctl = mdft::BATCH_CTL_HEADER + b->build.ncounters;
    count_phase_bytes(b.get());

    int per_cu = 0;
    mdft::hip_check(mdft::batch_occupancy(T, &per_cu), "occupancy");
    b->grid = std::max(1, ctx.cu_count() * std::max(per_cu, 1));

    const size_t njobs = b->build.jobs.size();
    const size_t jb = sizeof(BatchJob) * std::max<size_t>(njobs, 1);
    mdft::hip_check(hipMalloc(&b->jobs_dev, jb), "hipMalloc jobs");
    mdft::hip_check(hipHostMalloc(&b->jobs_pinned, jb, hipHostMallocDefault), "hipHostMalloc jobs");
    mdft::hip_check(hipMalloc(&b->items_dev, sizeof(uint2) * std::max<size_t>(b->nitems, 1)), "hipMalloc items");
    mdft::hip_check(hipMalloc(&b->ctl_dev, sizeof(uint32_t) * 2 * b->nctl), "hipMalloc ctl");
    if (b->nitems)
        mdft::hip_check(hipMemcpy(b->items_dev, b->build.items.data(), sizeof(uint2) * b->nitems,
                                  hipMemcpyHostToDevice),
                        "upload items");
    mdft::hip_check(hipMemset(b->ctl_dev, 0, sizeof(uint32_t) * 2 * b->nctl), "clear ctl");

    if (flags & MDFT_BATCH_GRAPH) {
        mdft::hip_check(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
        mdft::hip_check(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal), "begin capture");
        enqueue_batch(b.get(), b->cap_stream);
        mdft::hip_check(hipStreamEndCapture(b->cap_stream, &b->graph), "end capture");
        mdft::hip_check(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0), "graph instantiate");
    }
    *batch = b.release();
    MDFT_API_END
}

}  // extern "C"

namespace {

// Point the job table at src/dst (re-uploaded when they changed) and copy the
// 1x1 identities.
void prepare_exec(mdft_batch b, const void* src, void* dst, hipStream_t stream)
{
    if (src != b->cur_src || dst != b->cur_dst) {
        // the pinned staging copy of an earlier patch may still be in flight
        if (b->cur_src) mdft::hip_check(hipDeviceSynchronize(), "synchronize before re-patch");
        const float2* s = static_cast<const float2*>(src);
        float2* d = static_cast<float2*>(dst);
        const size_t njobs = b->build.jobs.size();
        for (size_t i = 0; i < njobs; ++i) {
            const mdft_desc& ds = b->descs[b->job_image[i]];
            BatchJob j = b->build.jobs[i];
            // a column job after a row job reads the row job's output
            const bool from_dst = b->job_is_col[i] && ds.width > 1;
            j.job.src = from_dst ? d + ds.dst_offset : s + ds.src_offset;
            j.job.dst = d + ds.dst_offset;
            b->jobs_pinned[i] = j;
        }
        if (njobs)
            mdft::hip_check(hipMemcpyAsync(b->jobs_dev, b->jobs_pinned, sizeof(BatchJob) * njobs,
                                           hipMemcpyHostToDevice, stream),
                            "upload jobs");
        b->cur_src = src;
        b->cur_dst = dst;
    }
    for (uint32_t i : b->copies) {
        const mdft_desc& ds = b->descs[i];
        if (static_cast<const float2*>(src) + ds.src_offset != static_cast<float2*>(dst) + ds.dst_offset)
            mdft::hip_check(hipMemcpyAsync(static_cast<float2*>(dst) + ds.dst_offset,
                                           static_cast<const float2*>(src) + ds.src_offset, sizeof(float2),
                                           hipMemcpyDeviceToDevice, stream),
                            "identity copy");
    }
}

}  // namespace

extern "C" {

int mdft_batch_exec_device(mdft_batch b, const void* src, void* dst, void* stream_p)
{
    MDFT_API_BEGIN
    if (!b) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!src || !dst) throw Error(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size");
    mdft::DeviceGuard g(b->device);
    hipStream_t stream = static_cast<hipStream_t>(stream_p);
    prepare_exec(b, src, dst, stream);
    if (b->gexec)
        mdft::hip_check(hipGraphLaunch(b->gexec, stream), "graph launch");
    else
        enqueue_batch(b, stream);
    b->last_stream = stream;
    MDFT_API_END
}

int mdft_batch_exec_phase(mdft_batch b, int phase, const void* src, void* dst, void* stream_p)
{
    MDFT_API_BEGIN
    if (!b) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!src || !dst) throw Error(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size");
    if (phase != 1 && phase != 2) throw Error(MDFT_ERR_INVALID, "mdft_batch_exec_phase: phase must be 1 or 2");
    if (b->flags & (MDFT_BATCH_PERSISTENT | MDFT_BATCH_DEPENDENT))
        throw Error(MDFT_ERR_STATE, "mdft_batch_exec_phase: only grid-shape batches run by phase");
    mdft::DeviceGuard g(b->device);
    hipStream_t stream = static_cast<hipStream_t>(stream_p);
    if (phase == 1) {
        prepare_exec(b, src, dst, stream);
        mdft::hip_check(mdft::launch_batch_grid(b->threads, b->jobs_dev, b->items_dev, b->nphase1, stream),
                        "batch launch (phase 1)");
    } else {
        if (src != b->cur_src || dst != b->cur_dst)
            throw Error(MDFT_ERR_STATE, "mdft_batch_exec_phase: phase 2 needs the pointers of the phase 1 call");
        mdft::hip_check(mdft::launch_batch_grid(b->threads_cols, b->jobs_dev, b->items_dev + b->nphase1,
                                                b->nitems - b->nphase1, stream),
                        "batch launch (phase 2)");
    }
    b->last_stream = stream;
    MDFT_API_END
}

int mdft_batch_info(mdft_batch b, uint32_t* items1, uint32_t* items2, uint64_t* bytes1, uint64_t* bytes2,
                    uint32_t* threads1, uint32_t* threads2)
{
    MDFT_API_BEGIN
    if (!b) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (items1) *items1 = b->nphase1;
    if (items2) *items2 = b->nitems - b->nphase1;
    if (bytes1) *bytes1 = b->phase_bytes[0];
    if (bytes2) *bytes2 = b->phase_bytes[1];
    if (threads1) *threads1 = (uint32_t)b->threads;
    if (threads2) *threads2 = (uint32_t)b->threads_cols;
    MDFT_API_END
}

int mdft_batch_check(mdft_batch b)
{
    MDFT_API_BEGIN
    if (!b) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    mdft::DeviceGuard g(b->device);
    mdft::hip_check(hipStreamSynchronize(b->last_stream), "synchronize");
    uint32_t e[2] = {0, 0};
    mdft::hip_check(hipMemcpy(&e[0], b->ctl_dev + 1, sizeof(uint32_t), hipMemcpyDeviceToHost), "read flag");
    mdft::hip_check(hipMemcpy(&e[1], b->ctl_dev + b->nctl + 1, sizeof(uint32_t), hipMemcpyDeviceToHost), "read flag");
    if (e[0] | e[1]) throw Error(MDFT_ERR_TIMEOUT, "batch dependency wait expired");
    MDFT_API_END
}

int mdft_batch_destroy(mdft_batch b)
{
    MDFT_API_BEGIN
    if (!b) return MDFT_OK;
    {
        mdft::DeviceGuard g(b->device);
        (void)hipDeviceSynchronize();
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
        if (b->jobs_dev) (void)hipFree(b->jobs_dev);
        if (b->jobs_pinned) (void)hipHostFree(b->jobs_pinned);
        if (b->items_dev) (void)hipFree(b->items_dev);
        if (b->ctl_dev) (void)hipFree(b->ctl_dev);
    }
    delete b;
    MDFT_API_END
}

}  // extern "C"
