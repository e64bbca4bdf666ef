// Construction of a batch's job table and work list, shared by the batch
// API and the server.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "runtime.hpp"

namespace mdft {

struct BatchBuild {
    int threads = 256;       // workgroup size of whole images and row tiles
    int threads_cols = 256;  // workgroup size of column tiles
    std::vector<BatchJob> jobs;
    std::vector<uint2> items;
    uint32_t nphase1 = 0;
    uint32_t ncounters = 0;
    std::vector<std::shared_ptr<const LinePlan>> keep;
};

// One transform of a batch: device pointers may be null and patched later
// through src_job/dst_job offsets (batch API) or given directly (server).
struct BatchEntry {
    uint32_t width, height;
    bool inverse;
    const float2* src;
    float2* dst;
};

// Workgroup size a set of sizes needs for whole images and row tiles: the
// smallest of 256, 512 and 1024 threads that holds one line of the longest
// padded length; throws MDFT_ERR_SIZE beyond the batch limit (8192).
int batch_threads_for(const uint32_t* widths, const uint32_t* heights, size_t count);

// Builds jobs and items.  dependent: one-launch ordering with per-image
// counters; otherwise two phases sorted by kernel variant.  job_image[i]
// receives the index of the entry each job belongs to and job_is_col[i]
// whether it is the column job.
// threads: tiles of whole images and rows; threads_cols: column tiles.
void build_batch(Context& ctx, const BatchEntry* entries, size_t count, int threads, int threads_cols,
                 bool dependent, BatchBuild& out, std::vector<uint32_t>* job_image = nullptr,
                 std::vector<uint8_t>* job_is_col = nullptr);

}  // namespace mdft
