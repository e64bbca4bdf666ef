// Text data files of the reference's DFT test (src/gpu/dft/test/test_data.h,
// test_data.cpp:36-125), header-only, so vectors can be exchanged with a
// reference build for direct diffing:
//
//   load_data<T>(path)                  "n1 n2" then n1*n2 lines "re im"
//   save_data<T>(path, x)               lines "re im" (no header)
//   generate_random_data<T>(path, n1, n2)
//
// Numbers are written in scientific notation with max_digits10 digits after
// the point and an explicit sign (test_data.cpp:86-90).  Random data come
// from a PCG-XSH-RR engine seeded through std::seed_seq like the reference's
// PCG (src/com/random/pcg.h:53-63, pcg.cpp:45-80) and the standard library's
// uniform_real_distribution(-1, 1), so a build with the same standard library
// writes the same file.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <ios>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace mdft::data {

class DataError final : public std::runtime_error {
public:
    explicit DataError(const std::string& s) : std::runtime_error(s) {}
};

// PCG-XSH-RR 64/32 (O'Neill 2014): state and odd increment from four 32-bit
// words of std::seed_seq({seed}).
class Pcg final {
    std::uint64_t state_;
    std::uint64_t inc_;

public:
    using result_type = std::uint32_t;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return 0xffffffffu; }

    explicit Pcg(result_type seed)
    {
        std::seed_seq seq({seed});
        std::array<std::uint32_t, 4> w;
        seq.generate(w.begin(), w.end());
        state_ = std::uint64_t(w[0]) | (std::uint64_t(w[1]) << 32);
        inc_ = std::uint64_t(w[2]) | (std::uint64_t(w[3]) << 32) | 1u;
    }

    result_type operator()()
    {
        const std::uint64_t x = state_;
        state_ = x * 6364136223846793005ull + inc_;
        const std::uint32_t s = static_cast<std::uint32_t>(((x >> 18) ^ x) >> 27);
        const std::uint32_t r = static_cast<std::uint32_t>(x >> 59);
        return (s >> r) | (s << ((32 - r) & 31));
    }
};

template <typename T>
struct LoadData final {
    int n1;
    int n2;
    std::vector<std::complex<T>> data;
};

namespace detail {
template <typename T>
void number_format(std::ostream& f)
{
    f << std::scientific << std::setprecision(std::numeric_limits<T>::max_digits10) << std::showpos << std::showpoint;
}
}  // namespace detail

// test_data.cpp:36-81
template <typename T>
LoadData<T> load_data(const std::filesystem::path& file_name)
{
    constexpr int MAX_DIMENSION_SIZE = 1'000'000'000;
    std::ifstream file(file_name);
    int n1, n2;
    file >> n1 >> n2;
    if (!file) throw DataError("Data dimensions read error");
    if (n1 < 1 || n2 < 1) throw DataError("Dimensions must be positive numbers");
    if (n1 > MAX_DIMENSION_SIZE || n2 > MAX_DIMENSION_SIZE) throw DataError("Dimensions are too big");
    const long long count = static_cast<long long>(n1) * n2;
    std::vector<std::complex<T>> data(count);
    for (long long i = 0; i < count; ++i) {
        T re, im;
        file >> re >> im;
        if (!file) throw DataError("Error reading number " + std::to_string(i));
        data[i] = std::complex<T>(re, im);
    }
    return {n1, n2, std::move(data)};
}

// test_data.cpp:83-97
template <typename T>
void save_data(const std::filesystem::path& file_name, const std::vector<std::complex<T>>& x)
{
    std::ofstream file(file_name);
    detail::number_format<T>(file);
    for (const std::complex<T>& c : x) file << c.real() << ' ' << c.imag() << '\n';
}

// test_data.cpp:99-125.  The reference passes (n1 << 32) + n2 to a
// constructor taking a 32-bit value, so the seed is n2.
template <typename T>
void generate_random_data(const std::filesystem::path& file_name, int n1, int n2)
{
    if (n1 < 1 || n2 < 1) throw DataError("Wrong size " + std::to_string(n1) + " " + std::to_string(n2));
    Pcg engine(static_cast<Pcg::result_type>((static_cast<unsigned long long>(n1) << 32) + n2));
    std::uniform_real_distribution<T> urd(-1, 1);
    std::ofstream file(file_name);
    file << n1 << ' ' << n2 << '\n';
    detail::number_format<T>(file);
    for (long long i = 0; i < static_cast<long long>(n1) * n2; ++i) {
        const T re = urd(engine);
        const T im = urd(engine);
        file << re << ' ' << im << '\n';
    }
}

}  // namespace mdft::data
