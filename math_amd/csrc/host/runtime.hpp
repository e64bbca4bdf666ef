// Host runtime of the 2-D DFT: per-device context (twiddle table, cache of
// per-length line plans), transform enqueueing and error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "../kernels/launch.hpp"

namespace mdft {

// Exception carrying an MDFT_ERR_* code; the C ABI turns it into a return
// code plus a thread-local message.
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

void hip_check(hipError_t e, const char* what);

// compute_m of src/gpu/dft/function.cpp:32-45; throws "Error size <n> for
// compute m" outside [1, 2^29].
int compute_m(long long n);
int ilog2(uint64_t v);
bool is_pow2(uint64_t v);
// FFT shape (fft_device.hpp) of a line of length n along an axis: log2(n)
// for a power of two, 3 * 2^k for n = 3 * 2^k in [192, 6144] (both direct
// FFTs of length n); for other lengths the Bluestein convolution length, the
// reference's compute_m(n) = 2^j, or 3 * 2^k >= 2n - 1 when that is shorter:
// along columns for 3 * 2^k in [192, 6144], along rows in [3072, 6144] (the
// shorter row tiles measured slower with the extra passes, the longer ones
// faster); lines longer than one tile keep compute_m.
constexpr int F3_ROW_MIN_K = 10;
int line_shape(long long n, bool cols);
// index of a shape in per-shape tables, 0 .. NSHAPES - 1
constexpr int NSHAPES = MAX_LOGM + 1 + (F3_MAX_K - F3_MIN_K + 1);
inline int shape_index(int shape) { return is_f3(shape) ? MAX_LOGM + 1 + (f3_k(shape) - F3_MIN_K) : shape; }

// Lines whose padded length is 2^three_factor_min_logm() or more split in
// three factors (2^29 unless MDFT_LARGE_SPLIT3_LOGM in [15, 29] says otherwise,
// read once per process: a testing knob that runs the three-factor steps on
// lengths the oracle checks in full).
int three_factor_min_logm();
// the factors (log2) of a padded length 2^logm above one tile; l3 = 0: two
void large_split(int logm, int* l1, int* l2, int* l3);
// scratch of one axis pass over `count` lines of padded length m: count * m
// elements above one tile, none otherwise
uint64_t large_scratch_elems(long long m, long long count);

// Device tables of one line length n.
struct LinePlan {
    int n = 0;
    int m = 1;         // FFT length: n for a direct line (2^k or 3 * 2^k), the Bluestein convolution length otherwise
    int logm = 0;      // log2(m) for power-of-two m (lines longer than one tile: the four-step split)
    int shape = 0;     // FFT shape of the tile kernels (line_shape(n, axis))
    int padded = 1;    // the reference's padded length compute_m(n) (reported by mdft_plan_info)
    bool blue = false;  // Bluestein (n neither 2^k nor a direct 3 * 2^k)
    float2* chirp = nullptr;  // n entries, Bluestein only
    float2* dtab = nullptr;   // m entries, Bluestein only
    // tile-resident Bluestein lengths: the direction tables (dir_tables_len)
    float2* dir = nullptr;
    // lines longer than one tile: M = 2^logm1 * 2^logm2 (four-step), or
    // 2^logm1 * 2^logm2 * 2^logm3 (logm3 > 0: three factors), with a
    // two-level table of exp(-2*pi*i*j/M)
    int logm1 = 0, logm2 = 0, logm3 = 0;
    float2* tw_lo = nullptr;
    float2* tw_hi = nullptr;
    uint64_t bytes = 0;
    bool large() const { return logm > MAX_LOGM; }
    // scratch (complex<float> elements) one axis pass over `count` lines needs
    uint64_t scratch_elems(int count) const { return large_scratch_elems(m, count); }
    ~LinePlan();
};

// Geometry and twiddles of a power-of-two long FFT of length m (a view: the
// tables belong to a LinePlan).
struct LargeFft {
    int64_t m = 0;
    int logm1 = 0, logm2 = 0, logm3 = 0;
    const float2* tw_lo = nullptr;
    const float2* tw_hi = nullptr;
};

// The first step's load and the last step's store of a three-factor FFT.
struct LargeEnds {
    int load_mode = 0;   // LargeLoad (plain, conj, chirp)
    int store_mode = 0;  // LargeStore of the last step (conj with scale)
    bool inverse = false;
    int64_t n = 0;       // line length (the chirp load's zero padding, the chirp store's crop)
    float scale = 1.0f;
    const float2* chirp = nullptr;
    const float2* dtab = nullptr;
};

class Context {
public:
    static Context& get(int device);
    // the device's context if one was created, else null (creates nothing)
    static Context* find(int device);
    static int current_device();
    static void set_thread_device(int device);

    int device() const { return device_; }
    int cu_count() const { return cu_count_; }
    const float2* twiddles(int shape) const { return twiddles_ + tw_off(shape); }

    // Cached, device-resident plan of one line length (thread-safe).  The
    // cache holds at most cache_limit() bytes of tables that no caller holds:
    // beyond it the least recently requested unused lengths are freed.
    std::shared_ptr<const LinePlan> line_plan(int n, bool cols);
    // the cached plan of length n, or null; builds and allocates nothing
    std::shared_ptr<const LinePlan> find_line_plan(int n, bool cols);
    uint64_t cached_bytes();
    uint32_t cached_lengths();
    uint64_t cache_limit();
    // set the limit and free unused tables beyond it at once
    void set_cache_limit(uint64_t bytes);

    // Enqueue one axis pass of one image (scratch: lines longer than one tile).
    void enqueue_axis(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls, int count,
                      bool inverse, bool cols, float2* scratch, hipStream_t stream) const;
    // Enqueue a whole w x h transform (rows then columns); scratch must hold
    // max(row.scratch_elems(h), col.scratch_elems(w)) elements.
    void enqueue_transform(const LinePlan& row, const LinePlan& col, int w, int h, bool inverse, const float2* src,
                           float2* dst, float2* scratch, hipStream_t stream) const;
    // Four-step FFT of `count` lines of a long length (forward FFT of src into
    // dst, both natural order), the building block of the long-line path.
    void enqueue_large_axis(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls,
                            int64_t count, bool inverse, float2* scratch, hipStream_t stream) const;
    // Three-factor power-of-two FFT of length f.m (f.logm3 > 0) of count
    // lines, src (element stride es, line stride ls) into dst (tes, tls),
    // through T (count * m elements); e gives the ends' load and store.
    void enqueue_large_fft3(const LargeFft& f, const float2* src, int64_t es, int64_t ls, float2* dst, int64_t tes,
                            int64_t tls, int64_t count, const LargeEnds& e, float2* T, hipStream_t stream) const;
    // its first two steps (A1 with e's load, A2), into T
    void enqueue_large_a1a2(const LargeFft& f, const float2* src, int64_t es, int64_t ls, int64_t count,
                            const LargeEnds& e, float2* T, hipStream_t stream) const;
    // a three-factor Bluestein line in five steps through T (count * m elements)
    void enqueue_large_blue3(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls,
                             int64_t count, bool inverse, float2* T, hipStream_t stream) const;
    // Forward power-of-two four-step FFT of length f.m, times scale
    // (scratch: count * m elements).
    void enqueue_large_fft(const LargeFft& f, const float2* src, float2* dst, int64_t es, int64_t ls, int64_t count,
                           float scale, float2* scratch, hipStream_t stream) const;

    AxisJob make_job(const LinePlan& lp, const float2* src, float2* dst, int64_t es, int64_t ls, int count,
                     bool inverse) const;

    // Tile-shape table of the single-image kernels: variant 0 = narrow
    // (256-thread) tiles, 1 = wide (1024-thread) tiles, per (FFT shape, blue, cols).
    int variant(int shape, bool blue, bool cols) const { return variant_[shape_index(shape)][blue][cols]; }
    // the table's choice, narrowed when wide tiles would leave CUs without work
    bool wide_tiles(int shape, bool blue, bool cols, int count) const;
    // workgroup size of a single-image pass of `count` lines (launch_axis)
    int tile_threads(int shape, bool blue, bool cols, int count) const;
    void set_variant(int shape, bool blue, bool cols, int v) { variant_[shape_index(shape)][blue][cols] = (uint8_t)v; }

private:
    explicit Context(int device);
    int device_;
    int cu_count_ = 0;
    float2* twiddles_ = nullptr;
    hipStream_t gen_stream_ = nullptr;
    double* scratch64_ = nullptr;
    size_t scratch64_doubles_ = 0;
    std::mutex mu_;
    struct CacheEntry {
        std::shared_ptr<const LinePlan> plan;
        uint64_t last_use = 0;
    };
    // keyed by length and FFT shape (a length has at most two: compute_m and 3 * 2^k)
    std::unordered_map<int64_t, CacheEntry> lines_;
    uint64_t use_clock_ = 0;
    uint64_t cache_limit_ = uint64_t(4) << 30;
    void trim_locked(uint64_t limit);
    uint8_t variant_[NSHAPES][2][2] = {};
};

// Static scratch of the allocation-free entry point (noalloc.cpp): its device
// address is looked up once per device when the device's context is created
// (the device is current then), and kept per device.
void noalloc_setup(int device);
float2* noalloc_scratch(int device);
uint64_t noalloc_scratch_elems();


// Orders the calls that share one scratch buffer (long-line plans) when a
// caller enqueues them on different streams: a call on another stream than
// the previous one first waits for the previous call's completion event.
// Calls recorded into a stream capture are not tracked (their replays share
// the buffer; a caller must not replay two of them concurrently).
class ScratchOrder {
public:
    void before(hipStream_t s);
    void after(hipStream_t s);
    void release();

private:
    bool capturing(hipStream_t s) const;
    hipEvent_t ev_ = nullptr;
    hipStream_t last_ = nullptr;
    bool pending_ = false;
};

// RAII guard that makes a device current for the calling thread.
class DeviceGuard {
public:
    explicit DeviceGuard(int device);
    ~DeviceGuard();

private:
    int prev_;
};

}  // namespace mdft
