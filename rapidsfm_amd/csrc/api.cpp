// api.cpp - error reporting, version and the host-side match filters.
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>
#include "common.hpp"

namespace rsfm_amd {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace rsfm_amd

extern "C" const char* rsfm_last_error(void) { return rsfm_amd::g_last_error.c_str(); }

extern "C" const char* rsfm_version(void) { return "rsfm_amd 0.1 (gfx950)"; }

// crossCheckMatches, bruteForceMatch.cpp:23-40: ascending query order,
// invalid forward indices skipped
extern "C" size_t rsfm_cross_check_host(const rsfm_best_match* fwd, uint32_t nbQueryDesc,
                                        const rsfm_best_match* bwd, uint32_t nbTrainDesc, rsfm_valid_match* out)
{
    size_t cnt = 0;
    for (uint32_t i = 0; i < nbQueryDesc; i++) {
        const uint32_t t = fwd[i].index;
        if (t == ~0u || t >= nbTrainDesc) continue;
        if (bwd[t].index == i) out[cnt++] = rsfm_valid_match{i, t, static_cast<float>(fwd[i].distance)};
    }
    return cnt;
}

// removeMatchConflicts, bruteForceMatch.cpp:42-66: per train descriptor keep
// the first query with the strictly smallest distance, ascending train order
extern "C" size_t rsfm_remove_conflicts_host(const rsfm_best_match* fwd, uint32_t nbQueryDesc,
                                             uint32_t nbTrainDesc, rsfm_valid_match* out)
{
    std::vector<rsfm_best_match> bwd(nbTrainDesc, rsfm_best_match{~0u, UINT32_MAX});
    for (uint32_t i = 0; i < nbQueryDesc; i++) {
        const rsfm_best_match& q = fwd[i];
        if (q.index == ~0u || q.index >= nbTrainDesc) continue;
        rsfm_best_match& t = bwd[q.index];
        if (q.distance < t.distance) t = rsfm_best_match{i, q.distance};
    }
    size_t cnt = 0;
    for (uint32_t j = 0; j < nbTrainDesc; j++)
        if (bwd[j].index != ~0u) out[cnt++] = rsfm_valid_match{bwd[j].index, j, static_cast<float>(bwd[j].distance)};
    return cnt;
}
