/*
 * rng.c - random streams used by the oracle and the synthetic generators.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * or_mt64_*: the published MT19937-64 generator (Matsumoto & Nishimura 2004),
 *   the engine behind std::mt19937_64 that ransac.hpp:35 (Picker::engine)
 *   draws sample indices from. Pinned by the C++ standard's known answer
 *   (10000th output of the default-seeded engine = 9981545732273789042) and
 *   against libstdc++'s std::mt19937_64 (tests/test_oracle_rng.py).
 * or_lcg_step: the glibc TYPE_0 LCG of RansacMatchFilter.cu:45-50. Pinned
 *   against glibc's own rand_r() (tests/test_oracle_rng.py).
 * or_mt64_normal: Box-Muller; pinned by moments and the closed form.
 */
#include "rsfm_oracle.h"
#include <math.h>

#define NN 312
#define MM 156
#define MATRIX_A 0xB5026F5AA96619E9ULL
#define UM 0xFFFFFFFF80000000ULL
#define LM 0x7FFFFFFFULL

void or_mt64_seed(or_mt64* s, uint64_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < NN; i++)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->mti = NN;
}

uint64_t or_mt64_next(or_mt64* s)
{
    static const uint64_t mag01[2] = {0ULL, MATRIX_A};
    uint64_t x;
    if (s->mti >= NN) {
        int i;
        for (i = 0; i < NN - MM; i++) {
            x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
            s->mt[i] = s->mt[i + MM] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        for (; i < NN - 1; i++) {
            x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
            s->mt[i] = s->mt[i + (MM - NN)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        x = (s->mt[NN - 1] & UM) | (s->mt[0] & LM);
        s->mt[NN - 1] = s->mt[MM - 1] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        s->mti = 0;
    }
    x = s->mt[s->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

double or_mt64_uniform(or_mt64* s)
{
    return (double)(or_mt64_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

double or_mt64_normal(or_mt64* s)
{
    const double u1 = or_mt64_uniform(s);
    const double u2 = or_mt64_uniform(s);
    return sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586476925 * u2);
}

int32_t or_lcg_step(int32_t* state)
{
    /* unsigned arithmetic: the reference's int32 overflow wraps the same way */
    const uint32_t next = ((uint32_t)*state * 1103515245u + 12345u) & 0x7fffffffu;
    *state = (int32_t)next;
    return (int32_t)next;
}
