// Serving loop: continuous batching of 2-D DFT requests under a fixed HBM
// budget (no counterpart in the reference, whose exec blocks per call,
// src/gpu/dft/compute.cpp:257-283).
//
// * The server owns an arena of `budget` bytes of device memory.  A request
//   needs a result slot of w*h*8 bytes; requests are admitted in arrival
//   order while their slot fits (head-of-line FIFO, so a large request is
//   not starved by small ones).  A request larger than the whole budget is
//   refused at submission.
// * A scheduler thread turns the admitted requests into micro-batches (the
//   batch job tables and work lists of batch_builder.cpp, one workgroup per
//   tile) and launches each on the next free stream, so requests that arrive
//   while others run join the next launch and overlap with them on the
//   device.  Completion is observed per micro-batch with an event, so a
//   request of kIsolatePoints or more runs in a micro-batch of its own: the
//   small requests behind it in the queue launch on the other streams instead
//   of completing with it.
// * The row pass reads the client's device input directly and writes the
//   slot; the column pass works in the slot.  The result stays in the slot
//   until the client releases it.  Line plans (chirp and diagonal tables) are
//   cached per length on the device and shared by every request.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_builder.hpp"
#include "capi_util.hpp"

using mdft::BatchJob;
using mdft::Context;
using mdft::Error;

namespace {

// requests of this many points or more launch alone (4 Mi points, 32 MiB of
// complex<float>: about 4 % of a stream with sides log-uniform in [32, 4096])
constexpr uint64_t kIsolatePoints = uint64_t(1) << 22;

using Clock = std::chrono::steady_clock;

double now_us()
{
    return std::chrono::duration<double, std::micro>(Clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device allocation, with coalescing.
class Arena {
public:
    void init(char* base, uint64_t size)
    {
        base_ = base;
        size_ = size;
        free_.clear();
        free_[0] = size;
        used_ = 0;
    }
    // returns false when no free range is large enough
    bool alloc(uint64_t bytes, uint64_t* off)
    {
        bytes = (bytes + 255) & ~uint64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const uint64_t rest = it->second - bytes, start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = (bytes + 255) & ~uint64_t(255);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* base() const { return base_; }
    uint64_t used() const { return used_; }
    uint64_t size() const { return size_; }

private:
    char* base_ = nullptr;
    uint64_t size_ = 0, used_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

enum ReqState { PENDING = 0, RUNNING = 1, DONE = 2 };

struct Request {
    uint32_t w = 0, h = 0;
    bool inverse = false;
    const float2* src = nullptr;
    hipEvent_t ready = nullptr;  // producer event of submit_stream (null: src already written)
    uint64_t bytes = 0, off = 0;
    int state = PENDING;
    double t_submit = 0, t_launch = 0, t_done = 0;
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<uint64_t> reqs;
    // one device table and its pinned staging copy: cap_jobs jobs, then items
    char* table_dev = nullptr;
    char* table_host = nullptr;
    size_t cap_jobs = 0, cap_items = 0;
};

}  // namespace

struct mdft_server_s {
    int device = 0;
    Arena arena;
    char* arena_mem = nullptr;
    uint32_t max_batch = 128;
    // batching window: while another micro-batch runs, a new one launches once
    // min_batch requests wait or the oldest has waited max_wait_us
    uint32_t min_batch = 1;
    double max_wait_us = 0.0;
    std::vector<Lane> lanes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    // line plans of every (length, FFT shape) the server has seen: resident
    // for its lifetime and counted against the budget with the result slots
    std::unordered_map<int64_t, std::shared_ptr<const mdft::LinePlan>> plans;
    uint64_t table_bytes = 0;
    std::vector<hipEvent_t> free_events;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread sched;
    // statistics
    uint64_t completed = 0, launches = 0, launched_reqs = 0;
    std::string sched_error;
};

namespace {

size_t table_bytes(size_t jobs, size_t items) { return sizeof(BatchJob) * jobs + sizeof(uint2) * items; }

uint64_t slot_bytes(uint64_t bytes) { return (bytes + 255) & ~uint64_t(255); }

int64_t plan_key(int n, int shape) { return (int64_t(shape) << 32) | uint32_t(n); }

// Makes the plans of a request's lines resident (built on the calling thread,
// not the scheduler's) and returns the table bytes they add.  Line i has
// length lengths[i] and runs along rows when i is even, along columns when
// it is odd (a length can have a different FFT shape on each axis).
uint64_t hold_plans(mdft_server_s* s, const uint32_t* lengths, size_t count,
                    std::vector<std::shared_ptr<const mdft::LinePlan>>& out)
{
    Context& ctx = Context::get(s->device);
    uint64_t add = 0;
    for (size_t i = 0; i < count; ++i) {
        const int n = (int)lengths[i];
        const bool cols = i & 1;
        const int64_t key = plan_key(n, mdft::line_shape(n, cols));
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->plans.count(key)) continue;
        }
        bool dup = false;
        for (auto& p : out) dup |= plan_key(p->n, p->shape) == key;
        if (dup) continue;
        out.push_back(ctx.line_plan(n, cols));
        add += out.back()->bytes;
    }
    return add;
}

void commit_plans(mdft_server_s* s, std::vector<std::shared_ptr<const mdft::LinePlan>>& held)
{
    for (auto& p : held)
        if (s->plans.emplace(plan_key(p->n, p->shape), p).second) s->table_bytes += p->bytes;
}

void grow(Lane& L, size_t njobs, size_t nitems)
{
    if (njobs <= L.cap_jobs && nitems <= L.cap_items) return;
    if (L.table_dev) mdft::hip_check(hipFree(L.table_dev), "hipFree");
    if (L.table_host) mdft::hip_check(hipHostFree(L.table_host), "hipHostFree");
    L.table_dev = L.table_host = nullptr;
    const size_t cj = std::max(njobs, 2 * L.cap_jobs), ci = std::max(nitems, 2 * L.cap_items);
    mdft::hip_check(hipMalloc(&L.table_dev, table_bytes(cj, ci)), "hipMalloc");
    mdft::hip_check(hipHostMalloc(&L.table_host, table_bytes(cj, ci), hipHostMallocDefault), "hipHostMalloc");
    L.cap_jobs = cj;
    L.cap_items = ci;
}

// Launch one micro-batch on lane L (scheduler thread, lock not held).
void launch_micro_batch(mdft_server_s* s, Lane& L, const std::vector<mdft::BatchEntry>& entries)
{
    Context& ctx = Context::get(s->device);
    std::vector<uint32_t> ws(entries.size()), hs(entries.size());
    for (size_t i = 0; i < entries.size(); ++i) {
        ws[i] = entries[i].width;
        hs[i] = entries[i].height;
    }
    const int T = mdft::batch_threads_for(ws.data(), hs.data(), entries.size());
    mdft::BatchBuild bb;
    mdft::build_batch(ctx, entries.data(), entries.size(), T, std::max(T, mdft::BATCH_COL_THREADS), false, bb);
    grow(L, bb.jobs.size(), bb.items.size());
    // jobs at the start of the table, items after the lane's job capacity: one upload
    BatchJob* jobs_host = reinterpret_cast<BatchJob*>(L.table_host);
    uint2* items_host = reinterpret_cast<uint2*>(L.table_host + sizeof(BatchJob) * L.cap_jobs);
    std::copy(bb.jobs.begin(), bb.jobs.end(), jobs_host);
    std::copy(bb.items.begin(), bb.items.end(), items_host);
    const BatchJob* jobs_dev = reinterpret_cast<const BatchJob*>(L.table_dev);
    const uint2* items_dev = reinterpret_cast<const uint2*>(L.table_dev + sizeof(BatchJob) * L.cap_jobs);
    if (!bb.items.empty()) {
        // the used part: [0, jobs) and [cap_jobs, cap_jobs + items) as one span
        const size_t span = sizeof(BatchJob) * L.cap_jobs + sizeof(uint2) * bb.items.size();
        mdft::hip_check(hipMemcpyAsync(L.table_dev, L.table_host, span, hipMemcpyHostToDevice, L.stream),
                        "upload tables");
    }
    // 1x1 requests are identities: copy the single value
    for (const auto& e : entries)
        if (e.width == 1 && e.height == 1)
            mdft::hip_check(hipMemcpyAsync(e.dst, e.src, sizeof(float2), hipMemcpyDeviceToDevice, L.stream), "copy");
    const uint32_t n = (uint32_t)bb.items.size();
    mdft::hip_check(mdft::launch_batch_grid(T, jobs_dev, items_dev, bb.nphase1, L.stream), "serve launch");
    mdft::hip_check(mdft::launch_batch_grid(bb.threads_cols, jobs_dev, items_dev + bb.nphase1, n - bb.nphase1,
                                            L.stream),
                    "serve launch");
    mdft::hip_check(hipEventRecord(L.done, L.stream), "event record");
}

// One tiny micro-batch before traffic (a whole 16x16 image in phase 1's
// 256-thread tiles, the 1000-point columns of a 2x1000 image in phase 2's
// 512-thread tiles): the first launches of the batch kernels in a process
// load their code and stalled the first requests of a stream by 17 ms.
void warm_up(mdft_server_s* s)
{
    const size_t need = sizeof(float2) * (16 * 16 + 2 * 1000);
    if (s->arena.size() < need) return;
    float2* a = reinterpret_cast<float2*>(s->arena_mem);  // in place on arena contents, results unused
    const std::vector<mdft::BatchEntry> entries = {{16, 16, false, a, a}, {2, 1000, false, a + 256, a + 256}};
    Lane& L = s->lanes[0];
    launch_micro_batch(s, L, entries);
    mdft::hip_check(hipStreamSynchronize(L.stream), "server warm-up");
}

void scheduler(mdft_server_s* s)
{
    try {
        mdft::DeviceGuard g(s->device);
        for (;;) {
            bool worked = false;
            // retire finished micro-batches
            for (Lane& L : s->lanes) {
                if (!L.busy) continue;
                const hipError_t q = hipEventQuery(L.done);
                if (q == hipErrorNotReady) continue;
                mdft::hip_check(q, "event query");
                const double t = now_us();
                std::lock_guard<std::mutex> lk(s->mu);
                for (uint64_t id : L.reqs) {
                    auto it = s->reqs.find(id);
                    if (it == s->reqs.end()) continue;
                    it->second.state = DONE;
                    it->second.t_done = t;
                    ++s->completed;
                }
                L.reqs.clear();
                L.busy = false;
                worked = true;
                s->cv_done.notify_all();
            }
            // admit and launch on a free lane
            Lane* free_lane = nullptr;
            for (Lane& L : s->lanes)
                if (!L.busy) {
                    free_lane = &L;
                    break;
                }
            std::vector<mdft::BatchEntry> entries;
            std::vector<hipEvent_t> waits;
            bool holding = false;  // requests wait for the batching window
            if (free_lane) {
                bool any_busy = false;
                for (Lane& L : s->lanes) any_busy |= L.busy;
                std::lock_guard<std::mutex> lk(s->mu);
                if (s->stop && s->pending.empty() && !any_busy) return;
                // an idle device launches at once; a busy one lets requests
                // accumulate into the next micro-batch for up to max_wait_us
                if (any_busy && !s->stop && !s->pending.empty() && s->pending.size() < s->min_batch &&
                    now_us() - s->reqs[s->pending.front()].t_submit < s->max_wait_us)
                    holding = true;
                while (!holding && !s->pending.empty() && entries.size() < s->max_batch) {
                    Request& r = s->reqs[s->pending.front()];
                    const bool alone = (uint64_t)r.w * r.h >= kIsolatePoints;
                    if (alone && !entries.empty()) break;  // the next free stream takes it
                    uint64_t off = 0;
                    // slots in use + this slot + resident tables within the budget
                    if (s->arena.used() + slot_bytes(r.bytes) + s->table_bytes > s->arena.size()) break;
                    if (!s->arena.alloc(r.bytes, &off)) break;  // head of line waits for memory
                    r.off = off;
                    r.state = RUNNING;
                    r.t_launch = now_us();
                    float2* slot = reinterpret_cast<float2*>(s->arena.base() + off);
                    entries.push_back({r.w, r.h, r.inverse, r.src, slot});
                    if (r.ready) {
                        waits.push_back(r.ready);
                        r.ready = nullptr;
                    }
                    free_lane->reqs.push_back(s->pending.front());
                    s->pending.pop_front();
                    if (alone) break;
                }
            }
            if (!entries.empty()) {
                // inputs produced on a client stream: the lane waits for them
                for (hipEvent_t e : waits) mdft::hip_check(hipStreamWaitEvent(free_lane->stream, e, 0), "wait input");
                if (!waits.empty()) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->free_events.insert(s->free_events.end(), waits.begin(), waits.end());
                }
                launch_micro_batch(s, *free_lane, entries);
                free_lane->busy = true;
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    ++s->launches;
                    s->launched_reqs += entries.size();
                }
                worked = true;
            }
            if (!worked) {
                bool any_busy = false;
                for (Lane& L : s->lanes) any_busy |= L.busy;
                std::unique_lock<std::mutex> lk(s->mu);
                if (s->stop && s->pending.empty() && !any_busy) return;
                if (any_busy || holding) {
                    lk.unlock();
                    std::this_thread::yield();  // poll the in-flight events
                } else {
                    s->cv_work.wait_for(lk, std::chrono::microseconds(200));
                }
            }
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->sched_error = e.what();
        s->stop = true;
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int mdft_server_create(mdft_server* server, uint64_t hbm_budget_bytes, uint32_t streams, uint32_t max_batch)
{
    MDFT_API_BEGIN
    if (!server) throw Error(MDFT_ERR_INVALID, "mdft_server_create: server is NULL");
    *server = nullptr;
    if (hbm_budget_bytes < 256) throw Error(MDFT_ERR_INVALID, "mdft_server_create: budget too small");
    Context& ctx = Context::get(Context::current_device());
    mdft::DeviceGuard g(ctx.device());
    auto s = std::make_unique<mdft_server_s>();
    s->device = ctx.device();
    s->max_batch = max_batch ? max_batch : 128;
    mdft::hip_check(hipMalloc(&s->arena_mem, hbm_budget_bytes), "hipMalloc server arena");
    s->arena.init(s->arena_mem, hbm_budget_bytes);
    s->lanes.resize(streams ? streams : 4);
    for (Lane& L : s->lanes) {
        mdft::hip_check(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking), "hipStreamCreate");
        mdft::hip_check(hipEventCreateWithFlags(&L.done, hipEventDisableTiming), "hipEventCreate");
        // sized up front for max_batch requests of up to 16384 tiles each, so
        // no launch has to reallocate (a hipFree would wait for the whole device)
        grow(L, 2 * (size_t)s->max_batch, std::max<size_t>(size_t(1) << 20, 16384 * (size_t)s->max_batch / 4));
    }
    warm_up(s.get());
    s->sched = std::thread(scheduler, s.get());
    *server = s.release();
    MDFT_API_END
}

int mdft_server_set_batching(mdft_server s, uint32_t min_batch, double max_wait_us)
{
    MDFT_API_BEGIN
    if (!s) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (min_batch < 1 || !(max_wait_us >= 0.0))
        throw Error(MDFT_ERR_INVALID, "mdft_server_set_batching: min_batch >= 1 and max_wait_us >= 0");
    std::lock_guard<std::mutex> lk(s->mu);
    s->min_batch = std::min(min_batch, s->max_batch);
    s->max_wait_us = max_wait_us;
    MDFT_API_END
}

int mdft_server_prepare(mdft_server s, const uint32_t* lengths, uint32_t count)
{
    MDFT_API_BEGIN
    if (!s) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!lengths && count) throw Error(MDFT_ERR_INVALID, "mdft_server_prepare: lengths is NULL");
    // each length along both axes
    std::vector<uint32_t> both;
    both.reserve(2 * (size_t)count);
    for (uint32_t i = 0; i < count; ++i) {
        mdft::compute_m(lengths[i]);  // validates the length
        both.push_back(lengths[i]);
        both.push_back(lengths[i]);
    }
    std::vector<std::shared_ptr<const mdft::LinePlan>> held;
    hold_plans(s, both.data(), both.size(), held);
    std::lock_guard<std::mutex> lk(s->mu);
    commit_plans(s, held);
    MDFT_API_END
}

static void server_submit(mdft_server s, uint32_t width, uint32_t height, int inverse, const void* src,
                          hipStream_t producer, bool ordered, uint64_t* ticket)
{
    if (!s) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!src || !ticket) throw Error(MDFT_ERR_INVALID, "mdft_server_submit: src and ticket must not be NULL");
    mdft::batch_threads_for(&width, &height, 1);  // size checks
    Request r;
    r.w = width;
    r.h = height;
    r.inverse = inverse != 0;
    r.src = static_cast<const float2*>(src);
    r.bytes = sizeof(float2) * (uint64_t)width * height;
    r.t_submit = now_us();
    const uint32_t lens[2] = {width, height};
    std::vector<std::shared_ptr<const mdft::LinePlan>> held;
    const uint64_t add = hold_plans(s, lens, 2, held);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->sched_error.empty()) throw Error(MDFT_ERR_DEVICE, "server stopped: " + s->sched_error);
    if (slot_bytes(r.bytes) + s->table_bytes + add > s->arena.size())
        throw Error(MDFT_ERR_BUDGET, "request of " + std::to_string(r.bytes) + " bytes (with " +
                                         std::to_string(s->table_bytes + add) +
                                         " bytes of resident tables) exceeds the HBM budget of " +
                                         std::to_string(s->arena.size()) + " bytes");
    commit_plans(s, held);
    if (ordered) {
        mdft::DeviceGuard g(s->device);
        hipEvent_t e = nullptr;
        if (!s->free_events.empty()) {
            e = s->free_events.back();
            s->free_events.pop_back();
        } else {
            mdft::hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        }
        const hipError_t rc = hipEventRecord(e, producer);
        if (rc != hipSuccess) {
            s->free_events.push_back(e);
            mdft::hip_check(rc, "hipEventRecord");
        }
        r.ready = e;
    }
    const uint64_t id = s->next_id++;
    s->reqs.emplace(id, r);
    s->pending.push_back(id);
    *ticket = id;
    s->cv_work.notify_one();
}

int mdft_server_submit(mdft_server s, uint32_t width, uint32_t height, int inverse, const void* src, uint64_t* ticket)
{
    MDFT_API_BEGIN
    server_submit(s, width, height, inverse, src, nullptr, false, ticket);
    MDFT_API_END
}

int mdft_server_submit_stream(mdft_server s, uint32_t width, uint32_t height, int inverse, const void* src,
                              void* stream, uint64_t* ticket)
{
    MDFT_API_BEGIN
    server_submit(s, width, height, inverse, src, static_cast<hipStream_t>(stream), true, ticket);
    MDFT_API_END
}

int mdft_server_wait(mdft_server s, uint64_t ticket, int64_t timeout_us, void** result, double* latency_us)
{
    MDFT_API_BEGIN
    if (!s) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) throw Error(MDFT_ERR_INVALID, "unknown or released ticket");
    auto ready = [&] { return it->second.state == DONE || !s->sched_error.empty(); };
    if (timeout_us < 0)
        s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), ready))
        throw Error(MDFT_ERR_TIMEOUT, "request not finished within the timeout");
    if (it->second.state != DONE) throw Error(MDFT_ERR_DEVICE, "server stopped: " + s->sched_error);
    if (result) *result = s->arena.base() + it->second.off;
    if (latency_us) *latency_us = it->second.t_done - it->second.t_submit;
    MDFT_API_END
}

int mdft_server_fetch(mdft_server s, uint64_t ticket, void* host_dst, size_t count)
{
    MDFT_API_BEGIN
    if (!s) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    const char* src = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) throw Error(MDFT_ERR_INVALID, "unknown or released ticket");
        if (it->second.state != DONE) throw Error(MDFT_ERR_STATE, "request still pending or running");
        if (!host_dst || count != (size_t)it->second.w * it->second.h)
            throw Error(MDFT_ERR_DATA_SIZE, "Wrong DFT buffer size");
        src = s->arena.base() + it->second.off;
    }
    mdft::DeviceGuard g(s->device);
    mdft::hip_check(hipMemcpy(host_dst, src, sizeof(float2) * count, hipMemcpyDeviceToHost), "fetch");
    MDFT_API_END
}

int mdft_server_release(mdft_server s, uint64_t ticket)
{
    MDFT_API_BEGIN
    if (!s) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) throw Error(MDFT_ERR_INVALID, "unknown or released ticket");
    if (it->second.state != DONE) throw Error(MDFT_ERR_STATE, "request still pending or running");
    s->arena.release(it->second.off, it->second.bytes);
    s->reqs.erase(it);
    s->cv_work.notify_one();
    MDFT_API_END
}

int mdft_server_stats(mdft_server s, mdft_server_counters* out)
{
    MDFT_API_BEGIN
    if (!s || !out) throw Error(MDFT_ERR_INVALID, "mdft_server_stats: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->completed = s->completed;
    out->pending = s->pending.size();
    out->in_flight = s->reqs.size() - s->pending.size();
    out->launches = s->launches;
    out->launched_requests = s->launched_reqs;
    out->budget_bytes = s->arena.size();
    out->used_bytes = s->arena.used();
    out->table_bytes = s->table_bytes;
    MDFT_API_END
}

int mdft_server_replay(mdft_server s, const mdft_request* reqs, uint32_t count, double* latency_us,
                       double* makespan_us)
{
    MDFT_API_BEGIN
    if (!s) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if ((!reqs || !latency_us) && count) throw Error(MDFT_ERR_INVALID, "mdft_server_replay: NULL argument");
    std::vector<uint64_t> tickets(count, 0);
    std::atomic<uint32_t> submitted{0};
    std::atomic<bool> failed{false};
    std::string err;
    std::mutex err_mu;
    // collector: waits for each request in order, records its latency, releases its slot
    std::thread collector([&] {
        for (uint32_t i = 0; i < count; ++i) {
            while (submitted.load(std::memory_order_acquire) <= i && !failed.load()) std::this_thread::yield();
            if (failed.load()) return;
            double lat = 0;
            int rc = mdft_server_wait(s, tickets[i], -1, nullptr, &lat);
            if (rc == MDFT_OK) rc = mdft_server_release(s, tickets[i]);
            if (rc != MDFT_OK) {
                std::lock_guard<std::mutex> lk(err_mu);
                err = mdft_last_error();
                failed = true;
                return;
            }
            latency_us[i] = lat;
        }
    });
    const double t0 = now_us();
    for (uint32_t i = 0; i < count && !failed.load(); ++i) {
        while (now_us() - t0 < reqs[i].arrival_us) {
        }
        int rc = mdft_server_submit(s, reqs[i].width, reqs[i].height, reqs[i].inverse, reqs[i].src, &tickets[i]);
        if (rc != MDFT_OK) {
            std::lock_guard<std::mutex> lk(err_mu);
            err = mdft_last_error();
            failed = true;
            break;
        }
        submitted.store(i + 1, std::memory_order_release);
    }
    collector.join();
    if (failed.load()) throw Error(MDFT_ERR_DEVICE, "replay failed: " + err);
    if (makespan_us) *makespan_us = now_us() - t0;
    MDFT_API_END
}

int mdft_server_destroy(mdft_server s)
{
    MDFT_API_BEGIN
    if (!s) return MDFT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->pending.clear();
        s->cv_work.notify_all();
    }
    if (s->sched.joinable()) s->sched.join();
    {
        mdft::DeviceGuard g(s->device);
        for (Lane& L : s->lanes) {
            if (L.stream) {
                (void)hipStreamSynchronize(L.stream);
                (void)hipStreamDestroy(L.stream);
            }
            if (L.done) (void)hipEventDestroy(L.done);
            if (L.table_dev) (void)hipFree(L.table_dev);
            if (L.table_host) (void)hipHostFree(L.table_host);
        }
        for (auto& kv : s->reqs)
            if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
        for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
        if (s->arena_mem) (void)hipFree(s->arena_mem);
    }
    delete s;
    MDFT_API_END
}

}  // extern "C"
