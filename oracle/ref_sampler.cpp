// Driver around the reference's header-only ransac.hpp (built from the
// reference tree by oracle/Makefile into oracle/_ref/). Test infrastructure
// only: it prints what the reference sampler and ransac<> loop produce so the
// oracle's restatement can be compared against them.
//   ref_sampler sample <n> <H> <seed>          -> H lines of 8 indices (TableSampler<u32,8>)
//   ref_sampler loop <n> <conf> <maxit> <seed> <modulus>
//                                             -> best score, draws, 8 best indices
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <stdexcept>
#include <type_traits>
#include <vector>
#include "ransac.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "sample" && argc == 5) {
        const uint32_t n = std::strtoul(argv[2], nullptr, 10);
        const uint32_t H = std::strtoul(argv[3], nullptr, 10);
        const uint32_t seed = std::strtoul(argv[4], nullptr, 10);
        NonDuplicateSampler::TableSampler<uint32_t, 8> s(n, seed);
        for (uint32_t h = 0; h < H; h++) {
            const auto idx = s();
            for (int i = 0; i < 8; i++) std::printf(i ? " %u" : "%u", idx[i]);
            std::printf("\n");
        }
        return 0;
    }
    if (mode == "loop" && argc == 7) {
        const uint32_t n = std::strtoul(argv[2], nullptr, 10);
        const float conf = std::strtof(argv[3], nullptr);
        const size_t maxit = std::strtoul(argv[4], nullptr, 10);
        const uint32_t seed = std::strtoul(argv[5], nullptr, 10);
        const uint64_t modulus = std::strtoull(argv[6], nullptr, 10);
        size_t calls = 0, best = 0;
        auto fn = [&](const std::array<uint32_t, 8>& idx) -> size_t {
            calls++;
            uint64_t s = 0;
            for (int i = 0; i < 8; i++) s = s * 1000003ULL + idx[i];
            const size_t r = (size_t)((s * 2654435761ULL) % modulus);
            return r;
        };
        const auto bi = ransac<decltype(fn), 8, uint32_t>(fn, n, conf, maxit, seed);
        best = fn(bi);
        calls--;
        std::printf("%zu %zu", best, calls);
        for (int i = 0; i < 8; i++) std::printf(" %u", bi[i]);
        std::printf("\n");
        return 0;
    }
    return 2;
}
