// Batched transforms of mixed sizes in one persistent launch.
//
// Each transform contributes a row job and a column job; each job is cut
// into tiles of whole lines.  The work list interleaves images, biggest
// first (rows of image 0, rows of image 1, columns of image 0, rows of image
// 2, ...), so a column job starts soon after its rows and finds them in L2 or
// the 256 MB MALL.  A line length gets the same tile shape in every batch,
// so a result is bitwise the same alone or in any mix.
#include "capi_util.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

using mdft::BatchJob;
using mdft::Context;
using mdft::Error;
using mdft::LinePlan;

struct mdft_batch_s {
    int device = 0;
    uint32_t flags = 0;
    std::vector<mdft_desc> descs;
    std::vector<std::shared_ptr<const LinePlan>> keep;
    // host copy of the jobs with offsets, patched with base pointers at exec
    std::vector<BatchJob> jobs;
    std::vector<uint64_t> job_src_off, job_dst_off;
    std::vector<int> job_src_is_dst;  // column job reading the row job's output
    std::vector<std::pair<uint64_t, uint64_t>> copies;  // 1x1 identities (src_off, dst_off)
    uint32_t nitems = 0, nrow_items = 0;
    uint32_t ncounters = 0;
    BatchJob* jobs_dev = nullptr;
    BatchJob* jobs_pinned = nullptr;
    uint2* items_dev = nullptr;
    uint32_t* ctl_dev = nullptr;  // two control blocks (two-phase uses both)
    uint32_t nctl = 0;
    int grid = 0;
    const void* cur_src = nullptr;
    void* cur_dst = nullptr;
    hipStream_t last_stream = nullptr;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t spin_limit = 1u << 26;
};

namespace {

void enqueue_batch(mdft_batch_s* b, hipStream_t stream)
{
    if (b->flags & MDFT_BATCH_TWO_PHASE) {
        if (b->nrow_items)
            mdft::hip_check(mdft::launch_batch(b->jobs_dev, b->items_dev, b->nrow_items, b->ctl_dev, b->nctl, b->grid,
                                               b->spin_limit, stream),
                            "batch launch (rows)");
        if (b->nitems > b->nrow_items)
            mdft::hip_check(mdft::launch_batch(b->jobs_dev, b->items_dev + b->nrow_items, b->nitems - b->nrow_items,
                                               b->ctl_dev + b->nctl, b->nctl, b->grid, b->spin_limit, stream),
                            "batch launch (columns)");
    } else if (b->nitems) {
        mdft::hip_check(mdft::launch_batch(b->jobs_dev, b->items_dev, b->nitems, b->ctl_dev, b->nctl, b->grid,
                                           b->spin_limit, stream),
                        "batch launch");
    }
}

}  // namespace

extern "C" {

int mdft_batch_create(mdft_batch* batch, const mdft_desc* descs, uint32_t count, uint32_t flags)
{
    MDFT_API_BEGIN
    if (!batch) throw Error(MDFT_ERR_INVALID, "mdft_batch_create: batch is NULL");
    *batch = nullptr;
    if (!descs && count) throw Error(MDFT_ERR_INVALID, "mdft_batch_create: descs is NULL");
    Context& ctx = Context::get(Context::current_device());
    mdft::DeviceGuard g(ctx.device());
    auto b = std::make_unique<mdft_batch_s>();
    b->device = ctx.device();
    b->flags = flags;
    b->descs.assign(descs, descs + count);
    const int maxlogm = mdft::batch_max_logm();

    struct Img {
        uint32_t idx;
        double work;
    };
    std::vector<Img> order;
    for (uint32_t i = 0; i < count; ++i) {
        const mdft_desc& d = descs[i];
        const int m1 = mdft::compute_m(d.width), m2 = mdft::compute_m(d.height);
        if (mdft::ilog2(m1) > maxlogm || mdft::ilog2(m2) > maxlogm)
            throw Error(MDFT_ERR_SIZE, "batch transform " + std::to_string(i) + " (" + std::to_string(d.width) + "x" +
                                           std::to_string(d.height) + ") exceeds the batch padded-length limit " +
                                           std::to_string(1 << maxlogm));
        order.push_back({i, (double)m1 * d.height + (double)m2 * d.width});
    }
    std::stable_sort(order.begin(), order.end(), [](const Img& a, const Img& c) { return a.work > c.work; });

    struct JobTiles {
        int job;
        int ntiles;
    };
    std::vector<JobTiles> rows, cols;  // per image in order (job -1 if absent)
    uint32_t counter = 0;
    for (const Img& im : order) {
        const mdft_desc& d = descs[im.idx];
        const int w = (int)d.width, h = (int)d.height;
        auto rp = ctx.line_plan(w);
        auto cp = ctx.line_plan(h);
        b->keep.push_back(rp);
        b->keep.push_back(cp);
        JobTiles r{-1, 0}, c{-1, 0};
        const int sig = (w > 1 && h > 1 && !(flags & MDFT_BATCH_TWO_PHASE)) ? (int)counter++ : -1;
        if (w > 1) {
            BatchJob j{};
            j.job = ctx.make_job(*rp, nullptr, nullptr, 1, w, h, d.inverse != 0);
            j.logm = rp->logm;
            j.blue = rp->blue;
            j.cols = 0;
            const int lpw = mdft::batch_lines_per_tile(rp->logm);
            j.ntiles = (h + lpw - 1) / lpw;
            j.dep = -1;
            j.dep_need = 0;
            j.sig = sig;
            r = {(int)b->jobs.size(), j.ntiles};
            b->jobs.push_back(j);
            b->job_src_off.push_back(d.src_offset);
            b->job_dst_off.push_back(d.dst_offset);
            b->job_src_is_dst.push_back(0);
        }
        if (h > 1) {
            BatchJob j{};
            j.job = ctx.make_job(*cp, nullptr, nullptr, w, 1, w, d.inverse != 0);
            j.logm = cp->logm;
            j.blue = cp->blue;
            j.cols = 1;
            const int lpw = mdft::batch_lines_per_tile(cp->logm);
            j.ntiles = (w + lpw - 1) / lpw;
            j.dep = (w > 1) ? sig : -1;
            j.dep_need = (w > 1) ? r.ntiles : 0;
            j.sig = -1;
            c = {(int)b->jobs.size(), j.ntiles};
            b->jobs.push_back(j);
            b->job_src_off.push_back(w > 1 ? d.dst_offset : d.src_offset);
            b->job_dst_off.push_back(d.dst_offset);
            b->job_src_is_dst.push_back(w > 1 ? 1 : 0);
        }
        if (w == 1 && h == 1) b->copies.push_back({d.src_offset, d.dst_offset});
        rows.push_back(r);
        cols.push_back(c);
    }

    // work list
    std::vector<uint2> items;
    auto add = [&](const JobTiles& jt) {
        for (int t = 0; t < jt.ntiles; ++t) items.push_back(make_uint2((uint32_t)jt.job, (uint32_t)t));
    };
    if (flags & MDFT_BATCH_TWO_PHASE) {
        for (auto& r : rows) add(r);
        b->nrow_items = (uint32_t)items.size();
        for (auto& c : cols) add(c);
    } else {
        // rows of k+1 before columns of k
        for (size_t k = 0; k < rows.size(); ++k) {
            add(rows[k]);
            if (k > 0) add(cols[k - 1]);
        }
        if (!cols.empty()) add(cols.back());
        b->nrow_items = 0;
    }
    b->nitems = (uint32_t)items.size();
    b->ncounters = counter;
    b->nctl = mdft::BATCH_CTL_HEADER + counter;

    int per_cu = 0;
    mdft::hip_check(mdft::batch_occupancy(&per_cu), "occupancy");
    if (per_cu < 1) per_cu = 1;
    b->grid = std::max(1, std::min<int>(ctx.cu_count() * per_cu, (int)std::max<uint32_t>(b->nitems, 1)));

    const size_t jb = sizeof(BatchJob) * std::max<size_t>(b->jobs.size(), 1);
    mdft::hip_check(hipMalloc(&b->jobs_dev, jb), "hipMalloc jobs");
    mdft::hip_check(hipHostMalloc(&b->jobs_pinned, jb, hipHostMallocDefault), "hipHostMalloc jobs");
    mdft::hip_check(hipMalloc(&b->items_dev, sizeof(uint2) * std::max<size_t>(items.size(), 1)), "hipMalloc items");
    mdft::hip_check(hipMalloc(&b->ctl_dev, sizeof(uint32_t) * 2 * b->nctl), "hipMalloc ctl");
    if (!items.empty())
        mdft::hip_check(hipMemcpy(b->items_dev, items.data(), sizeof(uint2) * items.size(), hipMemcpyHostToDevice),
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

int mdft_batch_exec_device(mdft_batch b, const void* src, void* dst, void* stream_p)
{
    MDFT_API_BEGIN
    if (!b) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!src || !dst) throw Error(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size");
    mdft::DeviceGuard g(b->device);
    hipStream_t stream = static_cast<hipStream_t>(stream_p);
    if (src != b->cur_src || dst != b->cur_dst) {
        // the pinned staging copy may still be in flight from an earlier patch
        if (b->last_stream || b->cur_src) mdft::hip_check(hipDeviceSynchronize(), "synchronize before re-patch");
        const float2* s = static_cast<const float2*>(src);
        float2* d = static_cast<float2*>(dst);
        for (size_t i = 0; i < b->jobs.size(); ++i) {
            BatchJob j = b->jobs[i];
            j.job.src = b->job_src_is_dst[i] ? d + b->job_src_off[i] : s + b->job_src_off[i];
            j.job.dst = d + b->job_dst_off[i];
            b->jobs_pinned[i] = j;
        }
        if (!b->jobs.empty())
            mdft::hip_check(hipMemcpyAsync(b->jobs_dev, b->jobs_pinned, sizeof(BatchJob) * b->jobs.size(),
                                           hipMemcpyHostToDevice, stream),
                            "upload jobs");
        b->cur_src = src;
        b->cur_dst = dst;
    }
    for (auto& c : b->copies)
        if (static_cast<const float2*>(src) + c.first != static_cast<float2*>(dst) + c.second)
            mdft::hip_check(hipMemcpyAsync(static_cast<float2*>(dst) + c.second,
                                           static_cast<const float2*>(src) + c.first, sizeof(float2),
                                           hipMemcpyDeviceToDevice, stream),
                            "identity copy");
    if (b->gexec)
        mdft::hip_check(hipGraphLaunch(b->gexec, stream), "graph launch");
    else
        enqueue_batch(b, stream);
    b->last_stream = stream;
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
