// Command-line wrapper of include/mdft_data.hpp for the tests:
//   data_tool gen <path> <n1> <n2> [f|d]   generate_random_data
//   data_tool pcg <seed> <count>           prints the count-th engine output
//   data_tool copy <in> <out>              load_data then save_data (float)
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mdft_data.hpp"

int main(int argc, char** argv)
{
    using namespace mdft::data;
    try {
        if (argc >= 5 && !std::strcmp(argv[1], "gen")) {
            const bool dbl = argc >= 6 && argv[5][0] == 'd';
            if (dbl)
                generate_random_data<double>(argv[2], std::atoi(argv[3]), std::atoi(argv[4]));
            else
                generate_random_data<float>(argv[2], std::atoi(argv[3]), std::atoi(argv[4]));
            return 0;
        }
        if (argc == 4 && !std::strcmp(argv[1], "pcg")) {
            Pcg e((Pcg::result_type)std::strtoul(argv[2], nullptr, 10));
            const long n = std::atol(argv[3]);
            Pcg::result_type v = 0;
            for (long i = 0; i < n; ++i) v = e();
            std::printf("%u\n", v);
            return 0;
        }
        if (argc == 4 && !std::strcmp(argv[1], "copy")) {
            const LoadData<float> d = load_data<float>(argv[2]);
            std::printf("%d %d\n", d.n1, d.n2);
            save_data<float>(argv[3], d.data);
            return 0;
        }
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 2;
    }
    std::fprintf(stderr, "usage: data_tool gen|pcg|copy ...\n");
    return 1;
}
