// Autotuner of the single-image tile shapes: times the narrow (256-thread)
// and wide (1024-thread) variant of every (FFT shape, Bluestein, axis) kernel
// on the current device and keeps the faster in the context's table.  The
// table can be saved to and loaded from a text file, one line per entry:
// "shape blue cols variant".  The tile size only groups lines into
// workgroups: a line's result is bitwise the same under either variant.
#include "capi_util.hpp"

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

using mdft::Context;
using mdft::Error;

namespace {

// average duration of `reps` enqueues of fn on stream s, after two warm-ups
template <class F>
float time_launches(hipStream_t s, F&& fn, int reps)
{
    hipEvent_t a, b;
    mdft::hip_check(hipEventCreate(&a), "hipEventCreate");
    mdft::hip_check(hipEventCreate(&b), "hipEventCreate");
    fn();
    fn();
    mdft::hip_check(hipEventRecord(a, s), "hipEventRecord");
    for (int i = 0; i < reps; ++i) fn();
    mdft::hip_check(hipEventRecord(b, s), "hipEventRecord");
    mdft::hip_check(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0;
    mdft::hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

int tile_threads(int shape, bool wide)
{
    const int tpl = mdft::m_of(shape) / mdft::ept_of(shape);
    const int base = wide ? 1024 : 256;
    return tpl > base ? tpl : base;
}

// every FFT shape of the tile kernels: powers of two, then 3 * 2^k
std::vector<int> all_shapes()
{
    std::vector<int> v;
    for (int logm = 1; logm <= mdft::MAX_LOGM; ++logm) v.push_back(logm);
    for (int k = mdft::F3_MIN_K; k <= mdft::F3_MAX_K; ++k) v.push_back(mdft::F3_BASE + k);
    return v;
}

bool valid_shape(int shape)
{
    return (shape >= 1 && shape <= mdft::MAX_LOGM) ||
           (mdft::is_f3(shape) && mdft::f3_k(shape) >= mdft::F3_MIN_K && mdft::f3_k(shape) <= mdft::F3_MAX_K);
}

}  // namespace

extern "C" {

int mdft_tune_auto(double* seconds)
{
    MDFT_API_BEGIN
    const auto t0 = std::chrono::steady_clock::now();
    Context& ctx = Context::get(Context::current_device());
    mdft::DeviceGuard g(ctx.device());
    hipStream_t s = nullptr;
    mdft::hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    constexpr size_t BUF_ELEMS = size_t(1) << 23;  // 64 MiB of complex<float>
    float2* buf = nullptr;
    mdft::hip_check(hipMalloc(&buf, sizeof(float2) * BUF_ELEMS), "hipMalloc tuning buffer");
    mdft::hip_check(hipMemsetAsync(buf, 0, sizeof(float2) * BUF_ELEMS, s), "hipMemsetAsync");
    try {
        for (int shape : all_shapes()) {
            if (tile_threads(shape, false) == tile_threads(shape, true)) continue;  // one tile size only
            const int m = mdft::m_of(shape);
            for (int blue = 0; blue < 2; ++blue) {
                if (blue && shape < 2) continue;
                // a length of this shape: m itself (direct), m/2 + 1 for a
                // power-of-two Bluestein convolution, m/2 - 1 for a 3 * 2^k one
                // (m/2 = 3 * 2^(k-1) may itself be a direct length)
                const int n = !blue ? m : mdft::is_f3(shape) ? m / 2 - 1 : m / 2 + 1;
                // 3 * 2^k convolutions are only chosen for columns; the tile
                // kernels of a shape are tuned along both axes all the same
                auto lp = ctx.line_plan(n, mdft::is_f3(shape));
                if (lp->shape != shape) throw Error(MDFT_ERR_STATE, "tune: length " + std::to_string(n) +
                                                                        " does not have the shape being tuned");
                const int count = (int)(BUF_ELEMS / (size_t)n);
                for (int cols = 0; cols < 2; ++cols) {
                    // rows: count lines of n; columns: an n-high image of count columns
                    const int64_t es = cols ? count : 1, ls = cols ? 1 : n;
                    const mdft::AxisJob j = ctx.make_job(*lp, buf, buf, es, ls, count, false);
                    float best = 0;
                    int best_v = 0;
                    for (int v = 0; v < 2; ++v) {
                        const int threads = v ? 1024 : 256;
                        auto launch = [&]() {
                            mdft::hip_check(mdft::launch_axis(shape, j, blue != 0, cols != 0, threads, s),
                                            "tuning launch");
                        };
                        const float ms = time_launches(s, launch, 5);
                        if (v == 0 || ms < best) {
                            best = ms;
                            best_v = v;
                        }
                    }
                    ctx.set_variant(shape, blue != 0, cols != 0, best_v);
                }
            }
        }
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(buf);
        (void)hipStreamDestroy(s);
        throw;
    }
    mdft::hip_check(hipStreamSynchronize(s), "tuning");
    (void)hipFree(buf);
    (void)hipStreamDestroy(s);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    MDFT_API_END
}

int mdft_tune_save(const char* path)
{
    MDFT_API_BEGIN
    if (!path) throw Error(MDFT_ERR_INVALID, "tune: path is NULL");
    Context& ctx = Context::get(Context::current_device());
    FILE* f = std::fopen(path, "w");
    if (!f) throw Error(MDFT_ERR_INVALID, std::string("tune: cannot write ") + path);
    std::fprintf(f, "# mdft tile shapes: shape blue cols variant (0 = 256-thread tiles, 1 = 1024-thread tiles;\n"
                    "# shape 1..14: 2^shape points, 38..43: 3 * 2^(shape - 32) points)\n");
    for (int shape : all_shapes())
        for (int blue = 0; blue < 2; ++blue)
            for (int cols = 0; cols < 2; ++cols)
                std::fprintf(f, "%d %d %d %d\n", shape, blue, cols, ctx.variant(shape, blue != 0, cols != 0));
    const bool ok = std::fclose(f) == 0;
    if (!ok) throw Error(MDFT_ERR_INVALID, std::string("tune: cannot write ") + path);
    MDFT_API_END
}

int mdft_tune_load(const char* path)
{
    MDFT_API_BEGIN
    if (!path) throw Error(MDFT_ERR_INVALID, "tune: path is NULL");
    FILE* f = std::fopen(path, "r");
    if (!f) throw Error(MDFT_ERR_INVALID, std::string("tune: cannot read ") + path);
    struct Entry {
        int logm, blue, cols, v;
    };
    std::vector<Entry> entries;
    char line[256];
    int lineno = 0;
    bool bad = false;
    while (std::fgets(line, sizeof line, f)) {
        ++lineno;
        if (line[0] == '#' || line[0] == '\n') continue;
        Entry e{};
        if (std::sscanf(line, "%d %d %d %d", &e.logm, &e.blue, &e.cols, &e.v) != 4 || !valid_shape(e.logm) ||
            (e.blue != 0 && e.blue != 1) || (e.cols != 0 && e.cols != 1) ||
            (e.v != 0 && e.v != 1)) {
            bad = true;
            break;
        }
        entries.push_back(e);
    }
    std::fclose(f);
    if (bad)
        throw Error(MDFT_ERR_INVALID,
                    std::string("tune: bad entry on line ") + std::to_string(lineno) + " of " + path);
    // validated as a whole before anything is applied
    Context& ctx = Context::get(Context::current_device());
    for (const Entry& e : entries) ctx.set_variant(e.logm, e.blue != 0, e.cols != 0, e.v);
    MDFT_API_END
}

}  // extern "C"
