// C++ test of the drop-in interface, following the flow of the reference's
// src/gpu/dft/test/test_dft.cpp: constant data (:383-407), random data at a
// small size and at 6007x997 (:409-462), re-planning at random sizes before
// the timed exec (:138-148), forward then inverse, and the comparison with a
// library result by relative L1 discrepancy <= 1e-4 (:59,63-93).  Here the
// library is the CPU oracle (test infrastructure): its FP32 restatement of
// the reference pipeline and its FP64 DFT; the FP64 comparison also applies
// the normalised-L2 bound 2e-6*log2(padded N).
#include <chrono>
#include <cmath>
#include <filesystem>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/mdft.hpp"
#include "../../include/mdft_data.hpp"
#include "../../oracle/oracle.h"

using Complex = std::complex<float>;

namespace {

int failures = 0;

void fail(const std::string& msg)
{
    std::printf("FAIL %s\n", msg.c_str());
    ++failures;
}

double tolerance(int w, int h)
{
    return 2e-6 * std::max(1.0, std::log2((double)orc_compute_m(w) * orc_compute_m(h)));
}

// test_dft.cpp:63-93: sum|a-b| / sum|a| with a the computed result
double discrepancy(const std::vector<Complex>& a, const std::vector<Complex>& b)
{
    double sum = 0, sum2 = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        sum += std::abs(a[i] - b[i]);
        sum2 += std::abs(a[i]);
    }
    return sum == 0 ? 0 : sum / sum2;
}

void compare(const std::string& name, int w, int h, bool inverse, const std::vector<Complex>& input,
             const std::vector<Complex>& got)
{
    // FP32 restatement of the reference pipeline, as the reference compares with FFTW
    std::vector<Complex> ref32(input);
    orc_dft2d_ref32(w, h, inverse ? 1 : 0, reinterpret_cast<float*>(ref32.data()), 65536, 8);
    const double d = discrepancy(got, ref32);
    std::printf("  %s %dx%d %s: discrepancy vs reference pipeline %.3e", name.c_str(), w, h,
                inverse ? "inverse" : "forward", d);
    if (d > 1e-4) fail(name + ": relative L1 above 1e-4");
    // FP64 DFT
    std::vector<double> ref64(2 * input.size());
    orc_dft2d_64_from32(w, h, inverse ? 1 : 0, reinterpret_cast<const float*>(input.data()), ref64.data(), 8);
    double m[4];
    orc_error_metrics(got.size(), reinterpret_cast<const float*>(got.data()), ref64.data(), m);
    std::printf(", normalised L2 vs FP64 %.3e (bound %.2e)\n", m[0], tolerance(w, h));
    if (m[0] > tolerance(w, h)) fail(name + ": normalised L2 above the bound");
}

std::vector<Complex> uniform_data(int w, int h, unsigned seed)
{
    std::vector<Complex> v((size_t)w * h);
    orc_fill(seed, 0, v.size(), reinterpret_cast<float*>(v.data()));
    return v;
}

// test_dft.cpp:122-156: re-plan at random sizes, then the timed plan + exec
void compute(mdft::ComputeVector* dft, bool inverse, int w, int h, std::vector<Complex>* data, unsigned seed)
{
    std::mt19937 engine(seed);
    std::uniform_int_distribution<int> uid(1, 3000);
    dft->create_buffers(uid(engine), uid(engine));
    dft->create_buffers(1, 1);
    dft->create_buffers(uid(engine), uid(engine));
    dft->create_buffers(1, uid(engine));
    dft->create_buffers(uid(engine), uid(engine));
    dft->create_buffers(uid(engine), 1);
    dft->create_buffers(uid(engine), uid(engine));
    const auto t0 = std::chrono::steady_clock::now();
    dft->create_buffers(w, h);
    dft->exec(inverse, data);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  HIP %s time: %.3f ms (plan + exec incl. copies)\n", inverse ? "inverse" : "forward", ms);
}

void run(const std::string& name, mdft::ComputeVector* dft, int w, int h, const std::vector<Complex>& source)
{
    std::vector<Complex> data(source);
    compute(dft, false, w, h, &data, 1);
    compare(name, w, h, false, source, data);
    const std::vector<Complex> forward(data);
    compute(dft, true, w, h, &data, 2);
    compare(name, w, h, true, forward, data);
    // the DFT-backend shape gives the same result
    std::vector<Complex> again(source);
    mdft::create_hip(w, h)->exec(false, &again);
    if (again != forward) fail(name + ": create_hip result differs from ComputeVector");
}

void constant_data_test(mdft::ComputeVector* dft)
{
    std::printf("----- Constant Data DFT Tests -----\n");
    const std::vector<Complex> source{{1, 0}, {2, 0}, {3, 0}, {4, 0}, {5, 0}, {6, 0}};
    const int n = (int)source.size() / 3;
    const int k = (int)source.size() / n;
    run("constant", dft, n, k, source);
    // closed form of the forward transform of this input
    std::vector<Complex> data(source);
    dft->create_buffers(n, k);
    dft->exec(false, &data);
    const float s3 = std::sqrt(3.0f);
    const std::vector<Complex> expect{{21, 0}, {-3, 0}, {-6, 2 * s3}, {0, 0}, {-6, -2 * s3}, {0, 0}};
    for (size_t i = 0; i < expect.size(); ++i)
        if (std::abs(data[i] - expect[i]) > 1e-5f) fail("constant: closed form");
}

// test_dft.cpp:409-430 and :222: the random input goes through the reference's
// text file format (generate, load), the forward output is saved the same way
void data_file_test(mdft::ComputeVector* dft, int n1, int n2)
{
    std::printf("----- Random Data File DFT Test -----\n");
    const auto dir = std::filesystem::temp_directory_path();
    const auto in = dir / ("mdft_dft_input_" + std::to_string(n1) + "x" + std::to_string(n2) + ".txt");
    const auto out = dir / ("mdft_dft_output_" + std::to_string(n1) + "x" + std::to_string(n2) + ".txt");
    mdft::data::generate_random_data<float>(in, n1, n2);
    const mdft::data::LoadData<float> d = mdft::data::load_data<float>(in);
    if (d.n1 != n1 || d.n2 != n2) fail("data file: dimensions");
    run("file", dft, d.n1, d.n2, d.data);
    std::vector<Complex> forward(d.data);
    dft->create_buffers(n1, n2);
    dft->exec(false, &forward);
    mdft::data::save_data(out, forward);
    std::vector<Complex> back(forward.size());
    {
        // the saved file re-read with a header holds exactly the computed values
        std::ifstream f(out);
        for (auto& c : back) {
            float re, im;
            f >> re >> im;
            c = {re, im};
        }
    }
    if (back != forward) fail("data file: saved output does not read back exactly");
    std::filesystem::remove(in);
    std::filesystem::remove(out);
}

// the batched power-of-two FFT (src/gpu/dft/fft.h:32-53) against the FP64 oracle, per line
void fft1d_test()
{
    std::printf("----- Batched 1-D FFT -----\n");
    const int count = 7, n = 512;
    std::vector<Complex> data = uniform_data(n, count, 4242);
    const std::vector<Complex> source(data);
    mdft::Fft fft(count, n);
    for (int inverse = 0; inverse < 2; ++inverse) {
        data = source;
        fft.run_for_data(inverse != 0, &data);
        double worst = 0;
        for (int j = 0; j < count; ++j) {
            std::vector<double> in(2 * n), ref(2 * n);
            for (int k = 0; k < n; ++k) {
                in[2 * k] = source[(size_t)j * n + k].real();
                in[2 * k + 1] = source[(size_t)j * n + k].imag();
            }
            orc_dft1d_64(n, inverse, in.data(), ref.data());
            double m[4];
            orc_error_metrics(n, reinterpret_cast<const float*>(data.data() + (size_t)j * n), ref.data(), m);
            worst = std::max(worst, m[0]);
        }
        std::printf("  %s: worst normalised L2 vs FP64 %.3e\n", inverse ? "inverse" : "forward", worst);
        if (worst > 2e-6 * std::log2((double)n)) fail("batched FFT: normalised L2 above the bound");
    }
}

void error_tests()
{
    std::printf("----- Error behaviour -----\n");
    auto expect_error = [](const char* what, auto&& fn) {
        try {
            fn();
            fail(std::string("no error for ") + what);
        } catch (const mdft::ErrorException& e) {
            if (std::string(e.what()).find(what) == std::string::npos)
                fail(std::string("expected '") + what + "', got '" + e.what() + "'");
            else
                std::printf("  %s: ok\n", what);
        }
    };
    auto dft = mdft::create_compute_vector();
    std::vector<Complex> v(6);
    expect_error("No DFT buffers", [&] { dft->exec(false, &v); });
    dft->create_buffers(2, 3);
    std::vector<Complex> w(5);
    expect_error("Wrong DFT buffer size", [&] { dft->exec(false, &w); });
    expect_error("Error size 0 for compute m", [&] { dft->create_buffers(0, 3); });
    expect_error("Error HIP data size", [&] { mdft::create_hip(0, 3); });
}

}  // namespace

int main(int argc, char** argv)
{
    const bool large = !(argc > 1 && std::string(argv[1]) == "--small-only");
    auto dft = mdft::create_compute_vector();
    constant_data_test(dft.get());
    std::printf("----- Random Data DFT Tests -----\n");
    std::mt19937 engine(20261202);
    std::uniform_int_distribution<int> uid(1, 100);
    for (int i = 0; i < 10; ++i) {
        const int w = uid(engine), h = uid(engine);
        run("random", dft.get(), w, h, uniform_data(w, h, (unsigned)(w * 1000 + h)));
    }
    if (large) run("random", dft.get(), 6007, 997, uniform_data(6007, 997, 997));
    data_file_test(dft.get(), 37, 23);
    fft1d_test();
    error_tests();
    std::printf(failures ? "DFT check FAILED (%d)\n" : "DFT check passed\n", failures);
    return failures ? 1 : 0;
}
