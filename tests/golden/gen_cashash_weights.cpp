// Writes the cascade-hash projection weights and bias base produced by the
// recipe of the reference's BruteForceMatcher constructor
// (bruteForceMatcher.cpp:127-141): one std::default_random_engine{11} stream,
// weights = int8(round(clamp(N(0, 32), -128, 127.01))) for 64 x 128 values,
// then biasBase[i] = int32(round(sum_j w[i][j] * clamp(N(0, 16), -64, 64))).
// std::normal_distribution is implementation-defined, so the output is tied
// to this toolchain's libstdc++; it is committed as a fixture.
// Output (binary, little-endian): 8192 x int8 weights, then 64 x int32 bias base.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

int main(int argc, char** argv)
{
    if (argc != 2) return 2;
    std::vector<int8_t> w(64 * 128);
    std::normal_distribution<float> wd{0.f, 32.f};
    std::default_random_engine rng{11u};
    std::generate(w.begin(), w.end(), [&]() { return int8_t(std::round(std::clamp(wd(rng), -128.f, 127.01f))); });
    std::vector<int32_t> b(64);
    std::normal_distribution<float> bd{0.f, 16.f};
    for (int i = 0; i < 64; i++) {
        float v = 0;
        for (int j = 0; j < 128; j++) v += w[i * 128 + j] * std::clamp(bd(rng), -64.f, 64.f);
        b[i] = int32_t(std::round(v));
    }
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 1;
    std::fwrite(w.data(), 1, w.size(), f);
    std::fwrite(b.data(), 4, b.size(), f);
    std::fclose(f);
    return 0;
}
