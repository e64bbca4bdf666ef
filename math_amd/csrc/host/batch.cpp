// Batched transforms of mixed sizes.
//
// Each image contributes a row job and a column job, cut into tiles of whole
// lines.  A small image whose rows and columns each fit one tile becomes a
// single "whole" item done by one workgroup.  Items are ordered by kernel
// variant (padded length, Bluestein or not, axis) so neighbouring workgroups
// run the same code.  A line length gets the same radix schedule in every
// batch, so a result is bitwise the same alone or in any mix.
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
    int threads = 256;
    std::vector<std::shared_ptr<const LinePlan>> keep;
    // host copy of the jobs with offsets, patched with base pointers at exec
    std::vector<BatchJob> jobs;
    std::vector<uint64_t> job_src_off, job_dst_off;
    std::vector<int> job_src_is_dst;                    // column job reading the row job's output
    std::vector<std::pair<uint64_t, uint64_t>> copies;  // 1x1 identities (src_off, dst_off)
    uint32_t nphase1 = 0, nitems = 0;
    uint32_t ncounters = 0;
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
};

namespace {

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
        mdft::hip_check(mdft::launch_batch_grid(T, b->jobs_dev, b->items_dev + b->nphase1, b->nitems - b->nphase1,
                                                stream),
                        "batch launch (phase 2)");
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
    const bool dependent = (flags & MDFT_BATCH_DEPENDENT) != 0;

    // validate sizes and pick the workgroup size before touching the device
    int maxlogm = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const int m1 = mdft::compute_m(descs[i].width), m2 = mdft::compute_m(descs[i].height);
        maxlogm = std::max({maxlogm, mdft::ilog2(m1), mdft::ilog2(m2)});
    }
    const int limit = mdft::batch_max_logm(512);
    if (maxlogm > limit) {
        for (uint32_t i = 0; i < count; ++i) {
            const mdft_desc& d = descs[i];
            if (mdft::ilog2(mdft::compute_m(d.width)) > limit || mdft::ilog2(mdft::compute_m(d.height)) > limit)
                throw Error(MDFT_ERR_SIZE, "batch transform " + std::to_string(i) + " (" + std::to_string(d.width) +
                                               "x" + std::to_string(d.height) +
                                               ") exceeds the batch padded-length limit " +
                                               std::to_string(1 << limit));
        }
    }
    const int T = maxlogm > mdft::batch_max_logm(256) ? 512 : 256;

    Context& ctx = Context::get(Context::current_device());
    mdft::DeviceGuard g(ctx.device());
    auto b = std::make_unique<mdft_batch_s>();
    b->device = ctx.device();
    b->flags = flags;
    b->threads = T;

    std::vector<uint32_t> order(count);
    std::iota(order.begin(), order.end(), 0u);
    auto work = [&](uint32_t i) {
        return (double)mdft::compute_m(descs[i].width) * descs[i].height +
               (double)mdft::compute_m(descs[i].height) * descs[i].width;
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return work(a) > work(c); });

    // item with its ordering key (variant first, then image order)
    struct Item {
        int phase;
        int logm, blue, kind;  // kind: 0 whole, 1 row tile, 2 column tile
        uint32_t seq;
        uint2 v;
    };
    std::vector<Item> items;
    uint32_t counter = 0, seq = 0;
    std::vector<std::pair<std::vector<uint2>, std::vector<uint2>>> dep_lists;  // per image: row items, col items
    for (uint32_t idx : order) {
        const mdft_desc& d = descs[idx];
        const int w = (int)d.width, h = (int)d.height;
        auto rp = ctx.line_plan(w);
        auto cp = ctx.line_plan(h);
        b->keep.push_back(rp);
        b->keep.push_back(cp);
        const int lpw_r = w > 1 ? mdft::batch_lines_per_tile(T, rp->logm) : 0;
        const int lpw_c = h > 1 ? mdft::batch_lines_per_tile(T, cp->logm) : 0;
        const bool whole = (w > 1 && h > 1 && h <= lpw_r && w <= lpw_c);
        const int ntr = w > 1 ? (h + lpw_r - 1) / lpw_r : 0;
        const int ntc = h > 1 ? (w + lpw_c - 1) / lpw_c : 0;
        const int sig = (dependent && w > 1 && h > 1 && !whole) ? (int)counter++ : -1;
        int rj = -1, cj = -1;
        if (w > 1) {
            BatchJob j{};
            j.job = ctx.make_job(*rp, nullptr, nullptr, 1, w, h, d.inverse != 0);
            j.logm = rp->logm;
            j.blue = rp->blue;
            j.cols = 0;
            j.ntiles = ntr;
            j.dep = -1;
            j.sig = sig;
            rj = (int)b->jobs.size();
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
            j.ntiles = ntc;
            j.dep = sig;
            j.dep_need = ntr;
            j.sig = -1;
            cj = (int)b->jobs.size();
            b->jobs.push_back(j);
            b->job_src_off.push_back(w > 1 ? d.dst_offset : d.src_offset);
            b->job_dst_off.push_back(d.dst_offset);
            b->job_src_is_dst.push_back(w > 1 ? 1 : 0);
        }
        if (w == 1 && h == 1) b->copies.push_back({d.src_offset, d.dst_offset});
        std::vector<uint2> ri, ci;
        if (whole) {
            items.push_back({1, rp->logm, rp->blue, 0, seq++, make_uint2(rj, mdft::BATCH_WHOLE | (uint32_t)cj)});
            ri.push_back(items.back().v);
        } else {
            for (int t = 0; t < ntr; ++t) {
                items.push_back({1, rp->logm, rp->blue, 1, seq++, make_uint2(rj, t)});
                ri.push_back(items.back().v);
            }
            for (int t = 0; t < ntc; ++t) {
                items.push_back({w > 1 ? 2 : 1, cp->logm, cp->blue, 2, seq++, make_uint2(cj, t)});
                ci.push_back(items.back().v);
            }
        }
        dep_lists.push_back({ri, ci});
    }

    std::vector<uint2> list;
    if (dependent) {
        // rows of image k+1 before columns of image k: a column tile only waits
        // on row tiles with smaller tickets
        for (size_t k = 0; k < dep_lists.size(); ++k) {
            list.insert(list.end(), dep_lists[k].first.begin(), dep_lists[k].first.end());
            if (k > 0) list.insert(list.end(), dep_lists[k - 1].second.begin(), dep_lists[k - 1].second.end());
        }
        if (!dep_lists.empty()) list.insert(list.end(), dep_lists.back().second.begin(), dep_lists.back().second.end());
        b->nphase1 = (uint32_t)list.size();
    } else {
        std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& c) {
            return std::make_tuple(a.phase, -a.logm, a.blue, a.kind, a.seq) <
                   std::make_tuple(c.phase, -c.logm, c.blue, c.kind, c.seq);
        });
        for (const Item& it : items) {
            if (it.phase == 1) ++b->nphase1;
            list.push_back(it.v);
        }
    }
    b->nitems = (uint32_t)list.size();
    b->ncounters = counter;
    b->nctl = mdft::BATCH_CTL_HEADER + counter;

    int per_cu = 0;
    mdft::hip_check(mdft::batch_occupancy(T, &per_cu), "occupancy");
    b->grid = std::max(1, ctx.cu_count() * std::max(per_cu, 1));

    const size_t jb = sizeof(BatchJob) * std::max<size_t>(b->jobs.size(), 1);
    mdft::hip_check(hipMalloc(&b->jobs_dev, jb), "hipMalloc jobs");
    mdft::hip_check(hipHostMalloc(&b->jobs_pinned, jb, hipHostMallocDefault), "hipHostMalloc jobs");
    mdft::hip_check(hipMalloc(&b->items_dev, sizeof(uint2) * std::max<size_t>(list.size(), 1)), "hipMalloc items");
    mdft::hip_check(hipMalloc(&b->ctl_dev, sizeof(uint32_t) * 2 * b->nctl), "hipMalloc ctl");
    if (!list.empty())
        mdft::hip_check(hipMemcpy(b->items_dev, list.data(), sizeof(uint2) * list.size(), hipMemcpyHostToDevice),
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
        // the pinned staging copy of an earlier patch may still be in flight
        if (b->cur_src) mdft::hip_check(hipDeviceSynchronize(), "synchronize before re-patch");
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
