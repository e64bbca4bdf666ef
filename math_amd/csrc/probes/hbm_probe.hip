// Memory-access probe for the tile kernels' I/O shapes on one GPU (gfx950).
//
// Copy kernels with the load/store pattern of an axis pass and no FFT work:
//  * the dispatch floor: an empty kernel launched back to back, and two
//    different empty kernels alternating (a row pass then a column pass);
//  * a 1024 x 1024 complex<float> image copied in the row-pass shape (one
//    64-thread workgroup per line, 16 points per thread at stride 64) and in
//    the column-pass shape (4 columns per 256-thread tile), alone and as a pair;
//  * column tiles of LPW columns (each row segment LPW * 8 bytes wide) over a
//    W x H image, with workgroups dealt to the XCDs round-robin (hardware
//    order) or remapped so that neighbouring tiles share an XCD, against
//    whole contiguous lines of the same size.
// Every kernel reads and writes each element once; rates are algorithmic
// bytes (8 read + 8 written per element) over event-timed launches.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                        \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

constexpr int EPT = 16;  // points per thread, as in the radix-16 tile kernels

__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int xcd = bid % 8, local = bid / 8, q = nblocks / 8, r = nblocks % 8;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

__global__ void k_empty(float2* p)
{
    if (blockIdx.x == 0x7fffffffu) p[threadIdx.x].x = 1.0f;  // never taken
}
__global__ void k_empty2(float2* p)
{
    if (blockIdx.x == 0x7fffffffu) p[threadIdx.x].y = 1.0f;  // never taken
}

// T/64 lines of 64*EPT points per workgroup, thread t holding t + i*64
template <int T>
__global__ __launch_bounds__(T) void k_rows(const float2* __restrict__ src, float2* __restrict__ dst, int len)
{
    const int line = blockIdx.x * (T / 64) + threadIdx.x / 64, t = threadIdx.x % 64;
    float2 v[EPT];
#pragma unroll
    for (int i = 0; i < EPT; ++i) v[i] = src[(size_t)line * len + t + i * 64];
#pragma unroll
    for (int i = 0; i < EPT; ++i) v[i].x += 1.0f;
#pragma unroll
    for (int i = 0; i < EPT; ++i) dst[(size_t)line * len + t + i * 64] = v[i];
}

// LPW columns of a W-wide image per workgroup, rows t + i*(T/LPW)
template <int T, int LPW, bool REMAP>
__global__ __launch_bounds__(T) void k_cols(const float2* __restrict__ src, float2* __restrict__ dst, int W)
{
    constexpr int TPL = T / LPW;
    const int ll = threadIdx.x % LPW, t = threadIdx.x / LPW;
    const int tile = REMAP ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int c = tile * LPW + ll;
    float2 v[EPT];
#pragma unroll
    for (int i = 0; i < EPT; ++i) v[i] = src[(size_t)(t + i * TPL) * W + c];
#pragma unroll
    for (int i = 0; i < EPT; ++i) v[i].x += 1.0f;
#pragma unroll
    for (int i = 0; i < EPT; ++i) dst[(size_t)(t + i * TPL) * W + c] = v[i];
}

// the same tiles over lines stored contiguously (line c at c*H)
template <int T, int LPW>
__global__ __launch_bounds__(T) void k_lines(const float2* __restrict__ src, float2* __restrict__ dst, int H)
{
    constexpr int TPL = T / LPW;
    const int ll = threadIdx.x / TPL, t = threadIdx.x % TPL;
    const int c = blockIdx.x * LPW + ll;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
        float2 x = src[(size_t)c * H + t + i * TPL];
        x.x += 1.0f;
        dst[(size_t)c * H + t + i * TPL] = x;
    }
}

template <class F>
float time_us(hipStream_t s, F f, int reps = 50)
{
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    for (int i = 0; i < 5; ++i) f();
    CK(hipEventRecord(a, s));
    for (int i = 0; i < reps; ++i) f();
    CK(hipEventRecord(b, s));
    CK(hipEventSynchronize(b));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, a, b));
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return ms * 1e3f / reps;
}

constexpr size_t MAX_ELEMS = size_t(8192) * 4096;

template <int T, int LPW>
void column_tiles(hipStream_t s, const float2* x, float2* y, int W)
{
    constexpr int TPL = T / LPW;
    const int H = TPL * EPT;
    // the grid covers the image exactly: W/LPW tiles of LPW columns, H rows
    if (W % LPW != 0 || size_t(W) * H > MAX_ELEMS) {
        std::printf("shape %d x %d does not fit the probe\n", W, H);
        std::exit(1);
    }
    const int grid = W / LPW;
    const double gb = 16.0 * W * H * 1e-3;  // bytes per microsecond -> GB/s / 1000 = TB/s
    const float hw = time_us(s, [&] { k_cols<T, LPW, false><<<grid, T, 0, s>>>(x, y, W); });
    const float rm = time_us(s, [&] { k_cols<T, LPW, true><<<grid, T, 0, s>>>(x, y, W); });
    const float ln = time_us(s, [&] { k_lines<T, LPW><<<grid, T, 0, s>>>(x, y, H); });
    std::printf("| %4d x %4d | %4d | %2d | %3d B | %7.1f us %5.2f TB/s | %7.1f us %5.2f TB/s | %7.1f us %5.2f TB/s |\n",
                W, H, T, LPW, LPW * 8, hw, gb / hw * 1e-3, rm, gb / rm * 1e-3, ln, gb / ln * 1e-3);
}

int main()
{
    hipStream_t s;
    CK(hipStreamCreate(&s));
    float2 *x, *y;
    CK(hipMalloc(&x, MAX_ELEMS * sizeof(float2)));
    CK(hipMalloc(&y, MAX_ELEMS * sizeof(float2)));
    CK(hipMemset(x, 0, MAX_ELEMS * sizeof(float2)));
    CK(hipMemset(y, 0, MAX_ELEMS * sizeof(float2)));

    std::printf("dispatch floor and a 1024 x 1024 image (50 back-to-back launches each)\n\n");
    std::printf("| kernels | us per launch group |\n|---|---|\n");
    std::printf("| empty, 1024 x 64 threads | %.2f |\n", time_us(s, [&] { k_empty<<<1024, 64, 0, s>>>(y); }));
    std::printf("| two different empty kernels | %.2f |\n",
                time_us(s, [&] { k_empty<<<1024, 64, 0, s>>>(y); k_empty2<<<256, 256, 0, s>>>(y); }));
    // 1024 lines of 1024 = 64 * EPT points: rows 1024 workgroups, columns 1024/4 tiles
    constexpr int N = 1024;
    std::printf("| row-pass copy (64 threads, one line each) | %.2f |\n",
                time_us(s, [&] { k_rows<64><<<N, 64, 0, s>>>(x, y, N); }));
    std::printf("| column-pass copy (256 threads, 4 columns, XCD remap) | %.2f |\n",
                time_us(s, [&] { k_cols<256, 4, true><<<N / 4, 256, 0, s>>>(y, y, N); }));
    std::printf("| row-pass copy then column-pass copy | %.2f |\n", time_us(s, [&] {
                    k_rows<64><<<N, 64, 0, s>>>(x, y, N);
                    k_cols<256, 4, true><<<N / 4, 256, 0, s>>>(y, y, N);
                }));

    std::printf("\ncolumn tiles of LPW columns (16 points per thread)\n\n");
    std::printf("| image | threads | LPW | row segment | hardware XCD order | XCD remap | contiguous lines |\n");
    std::printf("|---|---|---|---|---|---|---|\n");
    column_tiles<512, 2>(s, x, y, 2048);   // 4096 rows
    column_tiles<512, 2>(s, x, y, 8192);
    column_tiles<1024, 4>(s, x, y, 4096);  // 4096 rows
    column_tiles<512, 4>(s, x, y, 2048);   // 2048 rows
    column_tiles<1024, 8>(s, x, y, 4096);  // 2048 rows
    column_tiles<512, 8>(s, x, y, 4096);   // 1024 rows
    column_tiles<1024, 16>(s, x, y, 4096); // 1024 rows
    column_tiles<256, 16>(s, x, y, 4096);  // 256 rows
    CK(hipStreamSynchronize(s));
    CK(hipFree(x));
    CK(hipFree(y));
    CK(hipStreamDestroy(s));
    std::printf("\ndone\n");
    return 0;
}
