// Copies between pageable host memory (the caller's std::vector) and device
// memory.  Large transfers go through a ring of two pinned chunks: the CPU
// copies chunk k into pinned memory while the DMA engine moves chunk k-1
// (measured on a 33.5 MB image: a host-vector transform 18 % faster than
// with the runtime's pageable copies).  Below STAGE_MIN the runtime's own
// pageable path is faster (the event round trips of the ring cost more than
// they save), so those copies go to it directly.  Used by the host-vector
// exec calls (mdft_exec, mdft_fft_exec); one ring per plan, allocated on
// first use.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "runtime.hpp"

namespace mdft {

class HostStaging {
public:
    static constexpr size_t CHUNK = size_t(4) << 20;
    static constexpr size_t STAGE_MIN = size_t(16) << 20;

    HostStaging() = default;
    HostStaging(const HostStaging&) = delete;
    HostStaging& operator=(const HostStaging&) = delete;

    // host -> device on stream s; returns once every chunk has been handed to
    // the DMA engine (the last ones may still be in flight, ordered before
    // later work on s)
    void upload(void* dev, const void* host, size_t bytes, hipStream_t s)
    {
        if (bytes < STAGE_MIN) {
            hip_check(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s), "copy in");
            return;
        }
        ensure();
        size_t k = 0;
        for (size_t off = 0; off < bytes; off += CHUNK, ++k) {
            const size_t n = std::min(CHUNK, bytes - off);
            const int i = int(k & 1);
            // the slot's previous chunk (k - 2) must have left it
            if (k >= 2) hip_check(hipEventSynchronize(ev_[i]), "staging wait");
            std::memcpy(slot_[i], static_cast<const char*>(host) + off, n);
            hip_check(hipMemcpyAsync(static_cast<char*>(dev) + off, slot_[i], n, hipMemcpyHostToDevice, s),
                      "copy in");
            hip_check(hipEventRecord(ev_[i], s), "staging event");
        }
    }

    // device -> host after the work already on stream s; blocks until the
    // data is in host memory
    void download(void* host, const void* dev, size_t bytes, hipStream_t s)
    {
        if (bytes < STAGE_MIN) {
            hip_check(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s), "copy out");
            hip_check(hipStreamSynchronize(s), "copy out");
            return;
        }
        ensure();
        const size_t nchunks = (bytes + CHUNK - 1) / CHUNK;
        auto issue = [&](size_t k) {
            const size_t off = k * CHUNK, n = std::min(CHUNK, bytes - off);
            const int i = int(k & 1);
            hip_check(hipMemcpyAsync(slot_[i], static_cast<const char*>(dev) + off, n, hipMemcpyDeviceToHost, s),
                      "copy out");
            hip_check(hipEventRecord(ev_[i], s), "staging event");
        };
        for (size_t k = 0; k < std::min<size_t>(2, nchunks); ++k) issue(k);
        for (size_t k = 0; k < nchunks; ++k) {
            const size_t off = k * CHUNK, n = std::min(CHUNK, bytes - off);
            const int i = int(k & 1);
            hip_check(hipEventSynchronize(ev_[i]), "staging wait");
            std::memcpy(static_cast<char*>(host) + off, slot_[i], n);
            if (k + 2 < nchunks) issue(k + 2);
        }
    }

    // call with the owning device current and its streams idle
    void release()
    {
        for (int i = 0; i < 2; ++i) {
            if (slot_[i]) (void)hipHostFree(slot_[i]);
            if (ev_[i]) (void)hipEventDestroy(ev_[i]);
            slot_[i] = nullptr;
            ev_[i] = nullptr;
        }
    }

private:
    void ensure()
    {
        for (int i = 0; i < 2; ++i) {
            if (!slot_[i])
                hip_check(hipHostMalloc(reinterpret_cast<void**>(&slot_[i]), CHUNK, hipHostMallocDefault),
                          "hipHostMalloc staging");
            if (!ev_[i]) hip_check(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming), "hipEventCreate");
        }
    }

    char* slot_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
};

}  // namespace mdft
