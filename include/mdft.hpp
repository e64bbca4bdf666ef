// C++ interface of the 2-D DFT, header-only over the C ABI of mdft.h.
//
// Mirrors the reference's drop-in shapes:
//   * ComputeVector::create_buffers(width, height) / exec(inverse, &vector)
//     and create_compute_vector()      (src/gpu/dft/compute.h:52-69)
//   * DFT::exec(inverse, &vector) and create_hip(x, y), the HIP counterpart
//     of create_fftw / create_cufft    (src/dft/dft.h:25-31, fftw.h:28)
// Errors throw ErrorException with the reference's message text
// ("No DFT buffers", "Wrong DFT buffer size", "Error size <n> for compute
// m"), as ns::error does (src/com/error.h:26-42).
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "mdft.h"

namespace mdft {

class ErrorException final : public std::exception {
    std::string text_;
    int code_;

public:
    ErrorException(int code, std::string text) : text_(std::move(text)), code_(code) {}
    const char* what() const noexcept override { return text_.c_str(); }
    int code() const noexcept { return code_; }
};

inline void check(int rc)
{
    if (rc != MDFT_OK) throw ErrorException(rc, mdft_last_error());
}

// One width x height plan (owns its device tables and buffer).
class Plan {
    mdft_plan p_ = nullptr;

public:
    Plan() = default;
    Plan(unsigned width, unsigned height) { check(mdft_plan_create(&p_, width, height)); }
    ~Plan() { mdft_plan_destroy(p_); }
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    Plan(Plan&& o) noexcept : p_(std::exchange(o.p_, nullptr)) {}
    Plan& operator=(Plan&& o) noexcept
    {
        if (this != &o) {
            mdft_plan_destroy(p_);
            p_ = std::exchange(o.p_, nullptr);
        }
        return *this;
    }
    mdft_plan handle() const { return p_; }
    void exec(bool inverse, std::vector<std::complex<float>>* data)
    {
        check(mdft_exec(p_, inverse ? 1 : 0, data ? data->data() : nullptr, data ? data->size() : 0));
    }
    // Enqueue on a HIP stream (hipStream_t passed as void*); src may equal dst.
    void exec_device(bool inverse, const void* src, void* dst, void* stream = nullptr)
    {
        check(mdft_exec_device(p_, inverse ? 1 : 0, src, dst, stream));
    }
};

// src/gpu/dft/compute.h:52-60
class ComputeVector {
public:
    virtual ~ComputeVector() = default;
    virtual void create_buffers(unsigned width, unsigned height) = 0;
    virtual void exec(bool inverse, std::vector<std::complex<float>>* src) = 0;
    void exec(bool inverse, std::vector<std::complex<float>>& src) { exec(inverse, &src); }
};

namespace detail {
class ComputeVectorImpl final : public ComputeVector {
    Plan plan_;

public:
    void create_buffers(unsigned width, unsigned height) override
    {
        plan_ = Plan();  // the reference deletes the old buffers first (compute.cpp:229-233)
        plan_ = Plan(width, height);
    }
    void exec(bool inverse, std::vector<std::complex<float>>* src) override { plan_.exec(inverse, src); }
    using ComputeVector::exec;
};
}  // namespace detail

// src/gpu/dft/compute.h:68 (the device is the one selected with mdft_init, default 0)
inline std::unique_ptr<ComputeVector> create_compute_vector()
{
    return std::make_unique<detail::ComputeVectorImpl>();
}

// src/dft/dft.h:25-31
class DFT {
public:
    virtual ~DFT() = default;
    virtual void exec(bool inverse, std::vector<std::complex<float>>* data) = 0;
};

namespace detail {
class HipDft final : public DFT {
    Plan plan_;

public:
    HipDft(int x, int y) : plan_((unsigned)x, (unsigned)y) {}
    void exec(bool inverse, std::vector<std::complex<float>>* data) override { plan_.exec(inverse, data); }
};
}  // namespace detail

// The HIP counterpart of create_fftw(x, y) (src/dft/fftw.cpp:178-186).
inline std::unique_ptr<DFT> create_hip(int x, int y)
{
    if (!(x > 0 && y > 0)) throw ErrorException(MDFT_ERR_SIZE, "Error HIP data size");
    return std::make_unique<detail::HipDft>(x, y);
}

// count power-of-two FFTs of length n stored back to back, in place; the shape
// of ns::gpu::dft::Fft and create_fft (src/gpu/dft/fft.h:32-53) without the
// Vulkan handles.  Inverse is scaled by 1/n.
class Fft {
    mdft_fft f_ = nullptr;

public:
    Fft(unsigned count, unsigned n) { check(mdft_fft_create(&f_, count, n)); }
    ~Fft() { mdft_fft_destroy(f_); }
    Fft(const Fft&) = delete;
    Fft& operator=(const Fft&) = delete;
    void run_for_data(bool inverse, std::vector<std::complex<float>>* data)
    {
        check(mdft_fft_exec(f_, inverse ? 1 : 0, data ? data->data() : nullptr, data ? data->size() : 0));
    }
    // the analogue of Fft::commands(command_buffer, inverse): enqueue on a stream
    void commands(bool inverse, const void* src, void* dst, void* stream = nullptr) const
    {
        check(mdft_fft_exec_device(f_, inverse ? 1 : 0, src, dst, stream));
    }
};

inline std::unique_ptr<Fft> create_fft(unsigned count, unsigned n) { return std::make_unique<Fft>(count, n); }

// Many transforms of mixed sizes (mdft_batch_*).
class Batch {
    mdft_batch b_ = nullptr;

public:
    Batch(const std::vector<mdft_desc>& descs, uint32_t flags = MDFT_BATCH_GRAPH)
    {
        check(mdft_batch_create(&b_, descs.data(), (uint32_t)descs.size(), flags));
    }
    ~Batch() { mdft_batch_destroy(b_); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    void exec_device(const void* src, void* dst, void* stream = nullptr)
    {
        check(mdft_batch_exec_device(b_, src, dst, stream));
    }
    void check_errors() { check(mdft_batch_check(b_)); }
};

// Continuous-batching server (mdft_server_*).
class Server {
    mdft_server s_ = nullptr;

public:
    explicit Server(uint64_t budget_bytes, uint32_t streams = 0, uint32_t max_batch = 0)
    {
        check(mdft_server_create(&s_, budget_bytes, streams, max_batch));
    }
    ~Server() { mdft_server_destroy(s_); }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;
    uint64_t submit(uint32_t width, uint32_t height, bool inverse, const void* src)
    {
        uint64_t t = 0;
        check(mdft_server_submit(s_, width, height, inverse ? 1 : 0, src, &t));
        return t;
    }
    // returns the device address of the result; latency in microseconds
    void* wait(uint64_t ticket, double* latency_us = nullptr, int64_t timeout_us = -1)
    {
        void* r = nullptr;
        check(mdft_server_wait(s_, ticket, timeout_us, &r, latency_us));
        return r;
    }
    void fetch(uint64_t ticket, std::vector<std::complex<float>>* out)
    {
        check(mdft_server_fetch(s_, ticket, out->data(), out->size()));
    }
    void release(uint64_t ticket) { check(mdft_server_release(s_, ticket)); }
};

}  // namespace mdft
