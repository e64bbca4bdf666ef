// C entry points over the reference's own plan math, compiled from
// src/gpu/dft/function.cpp (+ src/com/error.cpp, src/com/log/write.cpp) where
// they lie.  TEST INFRASTRUCTURE ONLY: it pins the oracle's restatement of
// compute_m / compute_h / compute_h2 / shared_size / group_size
// (function.cpp:32-118); the product never loads it.
#include <src/gpu/dft/function.h>

#include <complex>
#include <cstring>
#include <exception>
#include <vector>

extern "C" {

// Returns -1 where the reference throws (its error() path).
int ref_compute_m(int n)
{
        try {
                return ns::gpu::dft::compute_m(n);
        } catch (const std::exception&) {
                return -1;
        }
}

void ref_compute_h(int n, int inverse, double coef, double* out)
{
        const std::vector<std::complex<double>> h = ns::gpu::dft::compute_h(n, inverse != 0, coef);
        std::memcpy(out, h.data(), sizeof(double) * 2 * h.size());
}

void ref_compute_h2(int n, int m, const double* h_in, double* out)
{
        std::vector<std::complex<double>> h(n);
        std::memcpy(h.data(), h_in, sizeof(double) * 2 * n);
        const std::vector<std::complex<double>> h2 = ns::gpu::dft::compute_h2(n, m, h);
        std::memcpy(out, h2.data(), sizeof(double) * 2 * h2.size());
}

unsigned ref_shared_size(unsigned dft_size, unsigned max_shared_memory_size)
{
        return ns::gpu::dft::shared_size<std::complex<float>>(dft_size, max_shared_memory_size);
}

unsigned ref_group_size(unsigned dft_size, unsigned max_group_size_x, unsigned max_group_invocations,
                        unsigned max_shared_memory_size)
{
        return ns::gpu::dft::group_size<std::complex<float>>(
                dft_size, max_group_size_x, max_group_invocations, max_shared_memory_size);
}
}
