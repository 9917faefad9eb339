// common.hpp - shared host/device helpers of the rsfm_amd library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include "rsfm_amd.h"

namespace rsfm_amd {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// per-thread last error message
void set_error(const std::string& msg);

inline int fail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

inline int check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(RSFM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return RSFM_OK;
}

inline int check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) return fail(RSFM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return RSFM_OK;
}

#define RSFM_RETURN_IF(cond, code, msg) \
    do {                                \
        if (cond) return ::rsfm_amd::fail(code, msg); \
    } while (0)

#define RSFM_CHECK(expr)                \
    do {                                \
        const int _rc = (expr);         \
        if (_rc != RSFM_OK) return _rc; \
    } while (0)

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// key of a BestMatch viewed as one little-endian u64: distance in the high
// word, index in the low word, so an unsigned minimum is the lexicographic
// (distance, index) minimum (the global key of bruteForceMatch.cuh:109-187)
__host__ __device__ inline uint64_t make_key(uint32_t dist, uint32_t idx)
{
    return (static_cast<uint64_t>(dist) << 32) | idx;
}
constexpr uint64_t kNoMatch = ~0ull;

// The dispatcher hands consecutive workgroups to the 8 XCDs round-robin, and
// each XCD has its own L2. Remapping the linear workgroup id so that runs of
// consecutive logical ids execute on one XCD keeps the workgroups that share
// an operand (all row blocks of one image pair) on the same L2.
constexpr uint32_t kNumXcd = 8;
struct Wg3 {
    uint32_t x, y, z;
};
__device__ inline Wg3 xcd_local_block()
{
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    const uint32_t total = gx * gy * gridDim.z;
    uint32_t L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (total % kNumXcd == 0) L = (L % kNumXcd) * (total / kNumXcd) + L / kNumXcd;
    return Wg3{L % gx, (L / gx) % gy, L / (gx * gy)};
}

}  // namespace rsfm_amd
