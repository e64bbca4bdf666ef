// Host-visible launch interface of the HIP kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"

namespace mdft {

// One axis pass of one image inside a batch.
struct BatchJob {
    AxisJob job;
    int shape;  // FFT shape of the job's lines (fft_device.hpp)
    int blue;
    int cols;
    int ntiles;
    int dep;       // counter to wait for (-1: none)
    int dep_need;  // value the counter must reach
    int sig;       // counter to bump after each tile (-1: none)
    int pad;
};

// ctl layout: [0] ticket, [1] error flag, [2..3] reserved, then counters
constexpr uint32_t BATCH_CTL_HEADER = 4;
// item.y flag: a whole small image, item.x = row job, item.y & ~flag = column job
constexpr uint32_t BATCH_WHOLE = 0x80000000u;

// ---- lines longer than one tile (four-step through a scratch buffer) ----
constexpr int LARGE_LO_BITS = 14;
constexpr int64_t LARGE_LO = int64_t(1) << LARGE_LO_BITS;
// padded lengths up to 2^28 split in two factors M1 * M2 (each <= 2^14), longer
// ones (2^29 and 2^30, the reference's longest: compute_m of n <= 2^29) in three
constexpr int LARGE_MAX_LOGM = 30;
// LARGE_LOAD_TWIDDLE: the load multiplies by the twiddle of LARGE_STORE_TWIDDLE
enum LargeLoad { LARGE_LOAD_PLAIN = 0, LARGE_LOAD_CONJ = 1, LARGE_LOAD_CHIRP = 2, LARGE_LOAD_TWIDDLE = 3 };
// LARGE_STORE_DIAG_FFT: conj(v * D) (natural index k2 + kmul*k1), a second FFT
// of the same length, then the twiddle of LARGE_STORE_TWIDDLE (the middle of a
// long Bluestein line in one step)
enum LargeStore { LARGE_STORE_PLAIN = 0, LARGE_STORE_TWIDDLE = 1, LARGE_STORE_DIAG = 2, LARGE_STORE_CHIRP = 3,
                  LARGE_STORE_CONJ = 4, LARGE_STORE_DIAG_FFT = 5 };

// One step of a long-line FFT: count sub-lines; sub-line L reads from
// src + (L / sub_in) * ls_in + r * ls2_in with r = L % sub_in and element
// stride es_in (with sub2_in > 0: (r % sub2_in) * ls2_in + (r / sub2_in) *
// ls3_in in place of r * ls2_in), and writes likewise with the *_out fields.
// The twiddle of LARGE_LOAD_TWIDDLE, LARGE_STORE_TWIDDLE and
// LARGE_STORE_DIAG_FFT is w_M^e of element pos, e = ((L % tmod) / tdiv) * tmul * pos.
// sub_*, sub2_*, tmod and tdiv are powers of two.
struct LargeJob {
    const float2* src;
    float2* dst;
    int64_t ls_in, ls2_in, es_in, sub_in;
    int64_t sub2_in, ls3_in;
    int64_t ls_out, ls2_out, es_out, sub_out;
    int64_t sub2_out, ls3_out;
    int64_t tmod, tdiv, tmul;
    int64_t count;
    int load_mode, store_mode, inverse, pad;
    int64_t n, m1, m2;
    int64_t kmul;  // natural index of the store modes DIAG, DIAG_FFT, CHIRP: q + kmul * pos
    // log2 of sub_in, sub2_in, sub_out, sub2_out, tmod, tdiv (all powers of two; -1: no second
    // split), filled in by launch_large_step
    float scale;
    const float2* chirp;
    const float2* dtab;
    const float2* tw_lo;
    const float2* tw_hi;
    const float2* tw;  // Stockham rows of the sub-FFT length
    int lg_sub_in, lg_sub2_in, lg_sub_out, lg_sub2_out, lg_tmod, lg_tdiv;
};

// min_blocks: smaller tiles when 1024-thread ones would give fewer workgroups
hipError_t launch_large_step(int logm, const LargeJob& job, int min_blocks, hipStream_t stream);
hipError_t gen_large_twiddles(float2* lo, float2* hi, int64_t m, hipStream_t stream);
hipError_t gen_h2_f32(float2* h2, const double2* chirp64, int64_t n, int64_t m, hipStream_t stream);

// A one-word launch from each kernel file (axis, image, long lines, batch): the
// first launch from a file loads all of its kernels (7.7 ms measured for the
// axis kernels), which the context does at creation rather than in a caller's
// first transform.
hipError_t warm_axis(uint32_t* word, hipStream_t stream);
hipError_t warm_image(uint32_t* word, hipStream_t stream);
hipError_t warm_large(uint32_t* word, hipStream_t stream);
hipError_t warm_batch(uint32_t* word, hipStream_t stream);

hipError_t gen_twiddles(float2* table, hipStream_t stream);
// Direction tables of a tile-resident Bluestein length (float2 entries):
// [load fwd c | load inv conj(c) | D | conj(D) | store fwd conj(c) | store inv c/n | m zeros]
// The zero tail lets a tile read the store blocks at every position of its
// live slots (up to M/2 + M/16 - 1 >= n) without a bounds check.
__host__ __device__ constexpr int64_t dir_tables_len(int64_t n, int64_t m) { return 4 * n + 3 * m; }
hipError_t gen_chirp(double2* chirp64, float2* chirp32, float2* dir, int n, int m, hipStream_t stream);
hipError_t gen_dir_diagonal(float2* dir, const float2* dtab, int n, int m, hipStream_t stream);
hipError_t gen_dtab(float2* dtab, const double2* chirp64, double* cos_scratch, int n, int m, hipStream_t stream);

// threads: 1024 wide tiles, 256 narrow, 128 or 64 small (clamped to one line per workgroup)
hipError_t launch_axis(int shape, const AxisJob& job, bool blue, bool cols, int threads, hipStream_t stream);

// image path: an axis pass with fused luminance input and/or magnitude output
// (mode = IO_LUM | IO_MAG bits), and the unfused conversions
hipError_t launch_image_axis(int shape, const AxisJob& job, const ImageIO& io, bool blue, bool cols, int mode,
                             bool wide, hipStream_t stream);
hipError_t launch_image_luminance(const ImageIO& io, float2* dst, int w, int h, hipStream_t stream);
hipError_t launch_image_magnitude(const ImageIO& io, const float2* src, int w, int h, hipStream_t stream);

// Workgroup size of batch column tiles: 2-16 columns per tile for padded
// lengths 1024-4096 (wider contiguous runs per row), two workgroups per CU.
constexpr int BATCH_COL_THREADS = 512;

// batch kernels: threads is 256 (padded lengths up to 4096), 512 or 1024 (up to 8192)
hipError_t launch_batch_grid(int threads, const BatchJob* jobs, const uint2* items, uint32_t nitems,
                             hipStream_t stream);
hipError_t launch_batch_persistent(int threads, const BatchJob* jobs, const uint2* items, uint32_t nitems,
                                   uint32_t* ctl, uint32_t nctl, int grid, uint32_t spin_limit, hipStream_t stream);
// lines of a shape one tile of `threads` threads holds (0: the shape does not fit)
int batch_lines_per_tile(int threads, int shape);
hipError_t batch_occupancy(int threads, int* blocks_per_cu);
int batch_max_logm(int threads);

}  // namespace mdft
