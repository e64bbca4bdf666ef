/*
 * mdft — MI355X-native 2-D DFT (drop-in for the 2-D DFT path of cppd/math).
 *
 * C ABI: plain pointers and sizes.  Data is row-major complex<float>
 * interleaved (re, im), element (x, y) at index y*width + x, width the fast
 * axis (src/gpu/dft/code/dft_mul.comp:94, src/dft/fftw.cpp:95).  Forward is
 * unnormalised, X[k] = sum x[j] exp(-2*pi*i*(j1*k1/w + j2*k2/h)); inverse uses
 * exp(+...) and is scaled by 1/(w*h) (src/gpu/dft/dft.cpp:326-347 with the
 * diagonals of dft.cpp:77-111; src/dft/fftw.cpp:140-150).  A size of 1 along
 * an axis leaves that axis unchanged (dft.cpp:330,339).
 *
 * Every entry point returns MDFT_OK (0) or a negative MDFT_ERR_* code; the
 * message of the last error on the calling thread is mdft_last_error().  The
 * message texts of the reference are kept where the reference has one.
 *
 * Threading: a plan or batch is used by one thread at a time (the reference's
 * Dft asserts thread affinity, dft.cpp:140,264,313,328); distinct plans may
 * be used concurrently.  The server's submit/poll/wait are thread-safe.
 */
#ifndef MDFT_H
#define MDFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDFT_OK 0
#define MDFT_ERR_INVALID (-1)     /* null pointer or bad argument */
#define MDFT_ERR_SIZE (-2)        /* "Error size <n> for compute m" (function.cpp:34-38) */
#define MDFT_ERR_NO_BUFFERS (-3)  /* "No DFT buffers" (compute.cpp:259-262) */
#define MDFT_ERR_DATA_SIZE (-4)   /* "Wrong DFT buffer size" (compute.cpp:264-267) */
#define MDFT_ERR_DEVICE (-5)      /* HIP runtime failure */
#define MDFT_ERR_BUDGET (-6)      /* request larger than the server's HBM budget */
#define MDFT_ERR_TIMEOUT (-7)     /* a bounded device-side wait expired */
#define MDFT_ERR_STATE (-8)       /* call not valid in the object's current state */

/* ---- library ------------------------------------------------------------ */

/* Message of the last error raised on this thread ("" if none). */
const char* mdft_last_error(void);

/* Library version string. */
const char* mdft_version(void);

/* Select the HIP device used by objects created afterwards on this thread
 * and build its twiddle tables.  Optional: device 0 is initialised lazily.
 * Replaces the reference's device choice of
 * create_compute_vector(DeviceSearchType) (src/gpu/dft/compute.h:68,
 * compute.cpp:286-305). */
int mdft_init(int device);

/* ---- one transform ------------------------------------------------------- */

typedef struct mdft_plan_s* mdft_plan;

/* Plan a width x height transform: padded sizes, FP64-generated chirp and
 * diagonal tables resident on the device.  Replaces
 * ComputeVector::create_buffers(width, height) (src/gpu/dft/compute.h:56,
 * compute.cpp:229-255) and Dft::create_buffers (dft.h:38, dft.cpp:262-309).
 * Every side the reference accepts runs (padded lengths up to 2^30).
 * Errors: MDFT_ERR_SIZE with "Error size <n> for compute m" when a side is
 * outside [1, 2^29]. */
int mdft_plan_create(mdft_plan* plan, uint32_t width, uint32_t height);

/* Free a plan and the device memory it owns (its image copy and long-line
 * scratch) (ComputeVector destructor / Dft::delete_buffers, dft.h:39,
 * dft.cpp:311-324).  The per-length chirp, diagonal and twiddle tables are
 * shared by every plan of the device through a table cache: they stay
 * resident while another plan, batch or server uses the length, and
 * afterwards until the cache exceeds its limit (mdft_cache_set_limit).  NULL
 * is accepted. */
int mdft_plan_destroy(mdft_plan plan);

/* Padded lengths the plan uses per axis: compute_m(side) (function.cpp:32-45),
 * which is the side itself for a power of two; and the plan's device bytes
 * (its share of the table cache counted in full, plus what it owns). */
int mdft_plan_info(mdft_plan plan, uint32_t* m1, uint32_t* m2, uint64_t* device_bytes);

/* ---- the per-device table cache ------------------------------------------- */

/* Bytes of device tables cached for this thread's device (shared twiddles
 * plus every cached length), the number of cached lengths, and the limit.
 * No counterpart in the reference, whose Dft owns its tables per plan
 * (src/gpu/dft/dft.cpp:262-309). */
int mdft_cache_info(uint64_t* bytes, uint32_t* lengths, uint64_t* limit);

/* Limit (bytes, default 4 GiB) of tables kept for lengths nothing uses: when
 * a new length pushes the cache above it, the least recently requested
 * unused lengths are freed.  Setting it frees the excess at once (0 frees
 * every unused length).  Freeing device memory waits for the device. */
int mdft_cache_set_limit(uint64_t bytes);

/* In-place transform of count complex<float> values in host memory, blocking
 * until the result is back.  Replaces ComputeVector::exec(inverse, &vector)
 * (src/gpu/dft/compute.h:58, compute.cpp:257-283) and ns::dft::DFT::exec
 * (src/dft/dft.h:29).  Errors: MDFT_ERR_NO_BUFFERS "No DFT buffers" for a
 * NULL plan; MDFT_ERR_DATA_SIZE "Wrong DFT buffer size" when data is NULL or
 * count != width*height. */
int mdft_exec(mdft_plan plan, int inverse, void* data, size_t count);

/* Enqueue the transform of a device image on a HIP stream (hipStream_t, NULL
 * for the default stream) without synchronising; src may equal dst.  The
 * analogue of recording Dft::compute_commands(command_buffer, inverse) into a
 * caller's command buffer (src/gpu/dft/dft.h:41, dft.cpp:326-347).  A plan
 * with a side whose padded length exceeds 16384 runs through a scratch buffer
 * it owns: a call on another stream than the plan's previous call waits on
 * the device for that call first (calls recorded into a stream capture are
 * not ordered this way: do not replay two such graphs of one plan at once). */
int mdft_exec_device(mdft_plan plan, int inverse, const void* src, void* dst, void* stream);

/* Allocation-free variant of mdft_exec_device for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc); no counterpart in the reference.  The
 * same kernels transform a width x height device image without a plan
 * object: the tables of the width as a row length and of the height as a
 * column length must already be resident (keep a plan of this size, or a
 * batch or server that holds a transform of this size, alive), and
 * lines whose padded length exceeds 16384 use a static per-device scratch of
 * 32 MiB (4 Mi complex<float>) instead of plan-owned memory.  Calls no
 * hipMalloc, hipFree or other allocating routine.  Errors: MDFT_ERR_SIZE when
 * a pass needs more scratch than that (message names the limit) or a side is
 * outside [1, 2^29]; MDFT_ERR_STATE when the device has no context yet
 * (mdft_init) or a length's tables are not resident.  Every call on a device
 * shares its scratch: calls that overlap on one GPU (other streams, replayed
 * graphs) race on it. */
int mdft_exec_device_noalloc(uint32_t width, uint32_t height, int inverse, const void* src, void* dst,
                             void* stream);

/* ---- tile-shape tuning ---------------------------------------------------- */

/* Tile shape the single-image kernels use for lines whose FFT has the given
 * shape: 1..14 for 2^shape points, 38..43 for 3 * 2^(shape - 32) points
 * (lines of exactly 3 * 2^k points, and the Bluestein convolution length used
 * when 3 * 2^k >= 2n - 1 is shorter than the reference's compute_m(n): along
 * columns from 192 points, along rows from 3072); Bluestein (blue != 0) or power-of-two, along
 * rows (cols == 0) or columns: 0 = 256-thread tiles, 1 = 1024-thread tiles
 * (four times the lines per tile).  No counterpart in the reference, whose
 * tile size follows the device's shared-memory limit
 * (src/gpu/dft/fft.cpp:51-61, function.cpp:94-118); this build picks it per
 * shape by measurement. */
int mdft_tune_get(int shape, int blue, int cols, int* variant);
int mdft_tune_set(int shape, int blue, int cols, int variant);

/* Autotuner: time both tile shapes of every (padded length, Bluestein, axis)
 * kernel on this thread's device (64 MiB of lines each, in a private buffer)
 * and keep the faster one in the table; *seconds (optional) receives the
 * time taken.  When the wide shape would leave CUs without a tile, the
 * narrow one runs regardless of the table. */
int mdft_tune_auto(double* seconds);

/* Save / load the table as text, one "shape blue cols variant" line per entry
 * ('#' starts a comment).  A file with a bad entry is rejected as a whole
 * (MDFT_ERR_INVALID). */
int mdft_tune_save(const char* path);
int mdft_tune_load(const char* path);

/* ---- batched power-of-two 1-D FFT ------------------------------------------ */

typedef struct mdft_fft_s* mdft_fft;

/* Plan `count` FFTs of power-of-two length n stored back to back (line j at
 * elements j*n .. j*n+n-1).  Forward unnormalised, inverse scaled by 1/n
 * (the reference divides by its tile length at the store and halves at every
 * global stage, src/gpu/dft/code/dft_fft_shared.comp:72,
 * dft_fft_global.comp:72-77).  n == 1 is a no-op plan.  Replaces
 * create_fft(device, family_indices, count, n) (src/gpu/dft/fft.h:49-53,
 * fft.cpp:293-330).  Errors (MDFT_ERR_SIZE, reference texts of
 * fft.cpp:298-306): "FFT size <n> is not positive", "FFT size <n> is not an
 * integral power of 2"; n above 2^29 ("... exceeds the longest supported
 * line (2^29)") or count*n above 2^40 are rejected. */
int mdft_fft_create(mdft_fft* fft, uint32_t count, uint32_t n);

/* In-place FFT of size complex<float> values in host memory, blocking.
 * Replaces Fft::run_for_data(inverse, data, ...) (fft.h:39-44,
 * fft.cpp:264-290).  Error MDFT_ERR_DATA_SIZE "FFT buffer size <size> must be
 * equal to data size <count*n>" (fft.cpp:276-280); a plan of n == 1 returns
 * without checking, as the reference does. */
int mdft_fft_exec(mdft_fft fft, int inverse, void* data, size_t size);

/* Enqueue the FFT of device memory on a stream (src may equal dst); the
 * analogue of Fft::commands(command_buffer, inverse) (fft.h:37,
 * fft.cpp:243-262).  n == 1 copies src to dst when they differ. */
int mdft_fft_exec_device(mdft_fft fft, int inverse, const void* src, void* dst, void* stream);

int mdft_fft_destroy(mdft_fft fft);

/* ---- the view's image path ------------------------------------------------- */

#define MDFT_PIXEL_RGBA32F 0u /* 4 x float per pixel (r, g, b, a) */
#define MDFT_PIXEL_RGBA8 1u   /* 4 x uint8 per pixel, UNORM (value / 255) */

typedef struct mdft_image_s* mdft_image;

/* Plan the image transform of a width x height rectangle: the luminance
 * 0.2126 r + 0.7152 g + 0.0722 b of each source pixel is the complex input
 * (src/gpu/dft/code/dft_copy_input.comp:37-54), the forward 2-D DFT runs, and
 * |X| / (width*height) is written to an R32F image of the rectangle's size
 * (dft_copy_output.comp:34-48, scale from compute.cpp:128-129).  Replaces
 * ComputeImage::create_buffers(sampler, input, output, rectangle, family)
 * (src/gpu/dft/compute.h:35-50, compute.cpp:103-134); the images are bound at
 * exec time.  Errors as mdft_plan_create, and MDFT_ERR_INVALID for an unknown
 * pixel format. */
int mdft_image_create(mdft_image* image, uint32_t width, uint32_t height, uint32_t pixel_format);

/* Enqueue the image transform on a stream.  src: device pixels of a
 * src_width x src_height image with a row pitch of src_pitch pixels; the
 * rectangle's origin is (x, y).  dst: device floats with a row pitch of
 * dst_pitch floats.  The luminance read is fused into the first axis pass and
 * the magnitude write into the last.  Replaces
 * ComputeImage::compute_commands(command_buffer) (compute.h:40,
 * compute.cpp:146-176).  MDFT_ERR_INVALID when the rectangle does not lie
 * inside the source or a pitch is smaller than its row (the reference asserts
 * these, compute.cpp:113-118). */
int mdft_image_exec_device(mdft_image image, const void* src, uint32_t src_width, uint32_t src_height,
                           uint64_t src_pitch, uint32_t x, uint32_t y, float* dst, uint64_t dst_pitch, void* stream);

/* Free the plan (ComputeImage::delete_buffers, compute.h:49); waits for the
 * device first.  NULL is accepted. */
int mdft_image_destroy(mdft_image image);

/* ---- many transforms of mixed sizes in one persistent launch ------------- */

typedef struct {
    uint32_t width;
    uint32_t height;
    int32_t inverse;
    uint32_t reserved;
    uint64_t src_offset; /* element (complex<float>) offsets into the base pointers */
    uint64_t dst_offset;
} mdft_desc;

typedef struct mdft_batch_s* mdft_batch;

#define MDFT_BATCH_GRAPH 1u       /* replay the launches as a captured hipGraph */
#define MDFT_BATCH_PERSISTENT 2u  /* persistent workgroups taking items from a ticket counter */
#define MDFT_BATCH_DEPENDENT 4u   /* one persistent launch; column tiles wait on per-image row counters */
#define MDFT_BATCH_TUNE 8u        /* grid shape: pick the column-tile size (256/512/1024 threads) by timing
                                     this batch's column phase on scratch images at create time */

/* Plan a batch of transforms (no counterpart in the reference, whose exec
 * handles one image per call).  Every side must run an FFT of at most 8192
 * points: powers of two up to 8192, 6144 (3 * 2^11, a direct FFT), other
 * sides up to 4097 (padded length compute_m(side) <= 8192). */
int mdft_batch_create(mdft_batch* batch, const mdft_desc* descs, uint32_t count, uint32_t flags);

/* Enqueue the whole batch on a stream.  Default shape: two launches (small
 * images whole plus row tiles, then column tiles), one workgroup per item;
 * the flags select persistent or dependency-counter shapes.  src may equal
 * dst (in place per transform).  A call whose src/dst differ from the
 * previous call's re-uploads the job table after synchronising the device;
 * make that call outside stream capture, after which calls with the same
 * pointers may be captured into a caller's graph. */
int mdft_batch_exec_device(mdft_batch batch, const void* src, void* dst, void* stream);

/* Enqueue one launch phase of a grid-shape batch (the default shape, with or
 * without MDFT_BATCH_GRAPH): phase 1 runs whole small images and row tiles,
 * phase 2 the column tiles and must follow a phase 1 call with the same
 * pointers.  For callers that interleave other work between the phases, and
 * for timing each phase.  MDFT_ERR_STATE for persistent shapes. */
int mdft_batch_exec_phase(mdft_batch batch, int phase, const void* src, void* dst, void* stream);

/* Work items of each phase, their algorithmic bytes (16 bytes per point per
 * axis pass: one read and one write of the image) and the workgroup sizes in
 * threads of each phase's tiles.  Any pointer may be NULL. */
int mdft_batch_info(mdft_batch batch, uint32_t* items1, uint32_t* items2, uint64_t* bytes1, uint64_t* bytes2,
                    uint32_t* threads1, uint32_t* threads2);

/* Device error flag of the last completed launch: 0, or MDFT_ERR_TIMEOUT if
 * a dependency wait expired.  Synchronises the batch's last stream. */
int mdft_batch_check(mdft_batch batch);

int mdft_batch_destroy(mdft_batch batch);

/* ---- serving: continuous batching under an HBM budget --------------------- */
/* No counterpart in the reference (its exec blocks per call,
 * src/gpu/dft/compute.cpp:257-283).  The server owns `hbm_budget_bytes` of
 * device memory; each request needs a result slot of width*height*8 bytes
 * and is admitted in arrival order when its slot, the slots in use and the
 * resident line-plan tables fit the budget.  A scheduler thread
 * launches the admitted requests as micro-batches on `streams` streams (at
 * most `max_batch` requests each; a request of 2^22 points or more runs in a
 * micro-batch of its own, so smaller ones do not complete with it), so
 * requests arriving while others run join the next launch and overlap with
 * them.  submit/wait/release/stats are thread-safe; destroy must not race
 * with them. */

typedef struct mdft_server_s* mdft_server;

typedef struct {
    uint64_t completed;
    uint64_t pending;           /* submitted, waiting for admission */
    uint64_t in_flight;         /* admitted and not yet released */
    uint64_t launches;          /* micro-batches launched */
    uint64_t launched_requests; /* requests over all micro-batches */
    uint64_t budget_bytes;
    uint64_t used_bytes;        /* slots currently held */
    uint64_t table_bytes;       /* resident line-plan tables (counted against the budget) */
} mdft_server_counters;

typedef struct {
    uint32_t width;
    uint32_t height;
    int32_t inverse;
    uint32_t reserved;
    const void* src;   /* device input, read when the request runs */
    double arrival_us; /* arrival time from the start of the replay */
} mdft_request;

/* streams = 0 and max_batch = 0 select the defaults (4 and 128).  Create runs
 * one tiny micro-batch on the arena (budgets of 18 KiB and more) so the batch
 * kernels are loaded before the first request. */
int mdft_server_create(mdft_server* server, uint64_t hbm_budget_bytes, uint32_t streams, uint32_t max_batch);

/* Batching window (no counterpart in the reference): while another
 * micro-batch is on the device, waiting requests launch together once
 * min_batch of them wait or the oldest has waited max_wait_us; an idle device
 * launches at once.  Defaults 1 and 0 (launch whenever a lane is free). */
int mdft_server_set_batching(mdft_server server, uint32_t min_batch, double max_wait_us);

/* Build the device-resident line plans of these lengths ahead of traffic. */
int mdft_server_prepare(mdft_server server, const uint32_t* lengths, uint32_t count);

/* Queue a request.  src is a device pointer to width*height complex<float>
 * that must be fully written when submit is called (work that produces it on
 * a stream must have completed: synchronise first, or use
 * mdft_server_submit_stream) and stay valid until the request completes.
 * The line plans of new lengths are built on the calling thread and stay
 * resident for the server's lifetime; their bytes count against the budget
 * together with the result slots.  MDFT_ERR_BUDGET when the slot plus the
 * resident tables exceed the whole budget; MDFT_ERR_SIZE beyond the batch
 * padded-length limit (8192). */
int mdft_server_submit(mdft_server server, uint32_t width, uint32_t height, int inverse, const void* src,
                       uint64_t* ticket);

/* As mdft_server_submit, for an input still being produced on `stream`
 * (hipStream_t, NULL for the default stream): the request's launch waits on
 * the device for the work enqueued on that stream before this call. */
int mdft_server_submit_stream(mdft_server server, uint32_t width, uint32_t height, int inverse, const void* src,
                              void* stream, uint64_t* ticket);

/* Wait for a request (timeout_us < 0: no limit; MDFT_ERR_TIMEOUT when it
 * expires).  result receives the device address of the output (valid until
 * release), latency_us the time from submit to the completion being seen. */
int mdft_server_wait(mdft_server server, uint64_t ticket, int64_t timeout_us, void** result, double* latency_us);

/* Copy a finished request's result (count complex<float>) to host memory. */
int mdft_server_fetch(mdft_server server, uint64_t ticket, void* host_dst, size_t count);

/* Return a finished request's slot to the budget. */
int mdft_server_release(mdft_server server, uint64_t ticket);

int mdft_server_stats(mdft_server server, mdft_server_counters* stats);

/* Replay a request trace: submits each request at its arrival time from one
 * thread while another waits for and releases them in order, recording each
 * latency (latency_us[i]) and the time from the first arrival to the last
 * completion. */
int mdft_server_replay(mdft_server server, const mdft_request* requests, uint32_t count, double* latency_us,
                       double* makespan_us);

int mdft_server_destroy(mdft_server server);

#ifdef __cplusplus
}
#endif

#endif /* MDFT_H */
