// rmf.hip - per-cell affine RANSAC match filter.
//
// Reference: rmf::cudaRansacMatchFilter (RansacMatchFilter.cu:394-419) and its
// kernels kernelFindBestAffineForCells (:196-279),
// kernelTryBestAffineOfOtherCellsAndVote (:281-338) and kernelComputeVotes
// (:341-391); host side RansacMatchFilter::getInlierMask (RansacMatchFilter.cpp:40-135).
// Semantics kept: 3x3-cell neighbourhoods, 3 distinct draws per hypothesis
// from the glibc LCG, Gauss elimination without pivoting, scale gate
// det in [1/16, 16], inlier radius threshold * min(sqrt(det), 1), strict
// best-count per thread, the CTA tree reduction of CtaReducer::findMax, and a
// vote for every inlier of a cell whose best affine has >= 6 inliers.
// Deliberate differences (determinism): the LCG seed is an argument instead of
// clock(); a cell with fewer than 9 neighbourhood matches stores a NaN affine
// (the reference leaves it unwritten); every thread votes with the CTA-best
// count (the reference tests partial per-thread counts before a barrier).
#include "common.hpp"
#include <algorithm>
#include <cmath>
#include <vector>

namespace rsfm_amd {
namespace rmf {

constexpr int kCta = 128;
constexpr uint32_t kMinInliers = 6;

struct Cells {
    uint32_t start[9], size[9], sum;
};

__device__ inline void cells_init(Cells& c, const uint32_t* bounds, uint32_t cols, uint32_t rows, uint32_t bx, uint32_t by)
{
    c.sum = 0;
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int m = (int)by - 1 + t / 3, n = (int)bx - 1 + t % 3;
        if (m >= 0 && m < (int)rows && n >= 0 && n < (int)cols) {
            const uint32_t idx = (uint32_t)m * cols + (uint32_t)n;
            c.start[t] = bounds[idx];
            c.size[t] = bounds[idx + 1] - bounds[idx];
        } else {
            c.start[t] = 0;
            c.size[t] = 0;
        }
        c.sum += c.size[t];
    }
}

__device__ inline uint32_t cells_map(const Cells& c, uint32_t idx)
{
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        if (c.size[i] == 0) continue;
        const uint32_t nxt = acc + c.size[i];
        if (idx < nxt) return c.start[i] + (idx - acc);
        acc = nxt;
    }
    return ~0u;
}

__device__ inline float sq_threshold(const float T[6], float threshold)
{
    const float sqrScale = T[0] * T[4] - T[1] * T[3];
    const float scale = sqrtf(sqrScale);
    const float s = (1.f < scale) ? 1.f : scale;  // std::min(scale, 1.f), NaN kept
    const float th = threshold * s;
    return th * th;
}

__device__ inline bool is_inlier(const float T[6], float4 p, float sq)
{
    const float d0 = ((T[0] * p.x + T[1] * p.y) + T[2]) - p.z;
    const float d1 = ((T[3] * p.x + T[4] * p.y) + T[5]) - p.w;
    return d0 * d0 + d1 * d1 < sq;
}

__device__ inline uint32_t count_inliers(const Cells& c, const float T[6], const float4* m, float threshold)
{
    const float sq = sq_threshold(T, threshold);
    uint32_t cnt = 0;
    for (int i = 0; i < 9; i++)
        for (uint32_t k = c.start[i]; k < c.start[i] + c.size[i]; k++) cnt += is_inlier(T, m[k], sq);
    return cnt;
}

// solveGaussElim semantics (SfmUtils.hpp:156-184): no pivoting, 3x3 with two right-hand sides
__device__ inline void solve_gauss3(float Ab[3][5], float X[3][2])
{
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float inv = 1.f / Ab[i][i];
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            const float factor = Ab[j][i] * inv;
#pragma unroll
            for (int k = i + 1; k < 5; k++) Ab[j][k] = Ab[j][k] - factor * Ab[i][k];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        X[r][0] = Ab[r][3];
        X[r][1] = Ab[r][4];
    }
#pragma unroll
    for (int i = 2; i >= 0; i--) {
        const float inv = 1.f / Ab[i][i];
        X[i][0] = X[i][0] * inv;
        X[i][1] = X[i][1] * inv;
#pragma unroll
        for (int j = i - 1; j >= 0; j--) {
            const float factor = Ab[j][i];
            X[j][0] = X[j][0] - factor * X[i][0];
            X[j][1] = X[j][1] - factor * X[i][1];
        }
    }
}

__device__ inline uint32_t lcg(uint32_t& s)
{
    s = (s * 1103515245u + 12345u) & 0x7fffffffu;
    return s;
}

struct Reducer {
    uint32_t cnt[kCta / 2];
    float aff[kCta / 2][6];
};

// CtaReducer::findMax: thread t keeps its value unless t + mid holds a strictly larger count
__device__ inline void cta_find_max(Reducer& R, uint32_t& cnt, float T[6])
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t mid = kCta / 2; mid != 0; mid /= 2) {
        if (tid >= mid && tid < 2 * mid) {
            R.cnt[tid - mid] = cnt;
#pragma unroll
            for (int q = 0; q < 6; q++) R.aff[tid - mid][q] = T[q];
        }
        __syncthreads();
        if (tid < mid && cnt < R.cnt[tid]) {
            cnt = R.cnt[tid];
#pragma unroll
            for (int q = 0; q < 6; q++) T[q] = R.aff[tid][q];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kCta) void best_affine_kernel(float* __restrict__ affine, const float4* __restrict__ m,
                                                           const uint32_t* __restrict__ bounds, uint32_t cols,
                                                           uint32_t rows, uint32_t nbTests, float threshold, uint32_t seed)
{
    Cells c;
    cells_init(c, bounds, cols, rows, blockIdx.x, blockIdx.y);
    const uint32_t rank = blockIdx.y * cols + blockIdx.x;
    if (c.sum < 9) {
        if (threadIdx.x < 6) affine[rank * 6 + threadIdx.x] = __builtin_nanf("");
        return;
    }
    __shared__ Reducer R;
    float best[6] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f};
    uint32_t bestCnt = 0;
    uint32_t st = seed + (kCta * rank + threadIdx.x) * 22695477u;
    for (uint32_t h = threadIdx.x; h < nbTests; h += kCta) {
        uint32_t idx[3];
        uint32_t got = 0;
        while (got < 3) {
            const uint32_t v = lcg(st) % c.sum;
            bool dup = false;
            for (uint32_t q = 0; q < got; q++) dup |= idx[q] == v;
            if (!dup) idx[got++] = v;
        }
        float Ab[3][5];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const float4 p = m[cells_map(c, idx[r])];
            Ab[r][0] = p.x;
            Ab[r][1] = p.y;
            Ab[r][2] = 1.f;
            Ab[r][3] = p.z;
            Ab[r][4] = p.w;
        }
        float X[3][2];
        solve_gauss3(Ab, X);
        const float T[6] = {X[0][0], X[1][0], X[2][0], X[0][1], X[1][1], X[2][1]};
        const float sqrScale = T[0] * T[4] - T[1] * T[3];
        if (sqrScale < 0.0625f || sqrScale > 16.f || !isfinite(sqrScale)) continue;
        const uint32_t k = count_inliers(c, T, m, threshold);
        if (k > bestCnt) {
            bestCnt = k;
#pragma unroll
            for (int q = 0; q < 6; q++) best[q] = T[q];
        }
    }
    cta_find_max(R, bestCnt, best);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 6; q++) affine[rank * 6 + q] = best[q];
    }
}

__device__ inline void vote(uint32_t* votesWords, uint32_t idx)
{
    atomicAdd(votesWords + (idx >> 2), 1u << (8 * (idx & 3)));
}

__global__ __launch_bounds__(kCta) void vote_kernel(uint32_t* __restrict__ votesWords, const float* __restrict__ affine,
                                                    const float4* __restrict__ m, const uint32_t* __restrict__ bounds,
                                                    uint32_t cols, uint32_t rows, float threshold, bool tryOther)
{
    Cells c;
    cells_init(c, bounds, cols, rows, blockIdx.x, blockIdx.y);
    const uint32_t rank = blockIdx.y * cols + blockIdx.x;
    const uint32_t nCells = cols * rows;
    __shared__ Reducer R;
    __shared__ float sBest[6];
    __shared__ uint32_t sBestCnt;
    float best[6];
#pragma unroll
    for (int q = 0; q < 6; q++) best[q] = affine[rank * 6 + q];
    uint32_t bestCnt = 0;
    if (tryOther) {
        for (uint32_t n = threadIdx.x; n < nCells; n += kCta) {
            float T[6];
#pragma unroll
            for (int q = 0; q < 6; q++) T[q] = affine[n * 6 + q];
            const uint32_t k = count_inliers(c, T, m, threshold);
            if (k > bestCnt) {
                bestCnt = k;
#pragma unroll
                for (int q = 0; q < 6; q++) best[q] = T[q];
            }
        }
        cta_find_max(R, bestCnt, best);
    } else if (threadIdx.x == 0) {
        bestCnt = count_inliers(c, best, m, threshold);
    }
    if (threadIdx.x == 0) {
        sBestCnt = bestCnt;
#pragma unroll
        for (int q = 0; q < 6; q++) sBest[q] = best[q];
    }
    __syncthreads();
    if (sBestCnt < kMinInliers) return;
    float T[6];
#pragma unroll
    for (int q = 0; q < 6; q++) T[q] = sBest[q];
    const float sq = sq_threshold(T, threshold);
    for (uint32_t k = threadIdx.x; k < c.sum; k += kCta) {
        const uint32_t mi = cells_map(c, k);
        if (is_inlier(T, m[mi], sq)) vote(votesWords, mi);
    }
}

__global__ void zero_words_kernel(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace rmf
}  // namespace rsfm_amd

using namespace rsfm_amd;

extern "C" int rsfm_rmf_filter(uint8_t* votes, float* affineScratch, const rsfm_pt_pair* matches, uint32_t nbMatches,
                               const uint32_t* bounds, uint32_t cols, uint32_t rows, uint32_t nbRansacTests,
                               float threshold, int tryOtherAffines, uint32_t seed, hipStream_t stream)
{
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(votes) % 4, RSFM_ERR_INVALID, "rmf: votes must be 4-byte aligned");
    RSFM_RETURN_IF(nbMatches > 0 && (votes == nullptr || matches == nullptr || bounds == nullptr || affineScratch == nullptr),
                   RSFM_ERR_INVALID, "rmf: null pointer");
    RSFM_RETURN_IF(cols == 0 || rows == 0, RSFM_ERR_INVALID, "rmf: cols and rows must be >= 1");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(matches) % 16, RSFM_ERR_INVALID, "rmf: matches must be 16-byte aligned");
    if (nbMatches == 0) return RSFM_OK;
    const uint32_t words = div_up(nbMatches, 4);
    rmf::zero_words_kernel<<<div_up(words, 256), 256, 0, stream>>>(reinterpret_cast<uint32_t*>(votes), words);
    RSFM_CHECK(check_launch("rmf zero votes"));
    if (nbMatches < 16) return RSFM_OK;
    const dim3 grid(cols, rows);
    const float4* m = reinterpret_cast<const float4*>(matches);
    rmf::best_affine_kernel<<<grid, rmf::kCta, 0, stream>>>(affineScratch, m, bounds, cols, rows, nbRansacTests,
                                                            threshold, seed);
    RSFM_CHECK(check_launch("rmf best affine"));
    rmf::vote_kernel<<<grid, rmf::kCta, 0, stream>>>(reinterpret_cast<uint32_t*>(votes), affineScratch, m, bounds, cols,
                                                     rows, threshold, tryOtherAffines != 0);
    return check_launch("rmf vote");
}

// host binning of RansacMatchFilter::getInlierMask (RansacMatchFilter.cpp:42-67)
extern "C" int rsfm_rmf_bin_host(int imgWidth, int imgHeight, uint32_t cellCols, const rsfm_pt_pair* matches,
                                 uint32_t nbMatches, uint32_t* cols, uint32_t* rows, uint32_t* cellWidth,
                                 rsfm_pt_pair* reordered, uint32_t* bounds, uint32_t* order)
{
    RSFM_RETURN_IF(imgWidth <= 0 || imgHeight <= 0 || cellCols == 0, RSFM_ERR_INVALID, "rmf bin: bad image size");
    const uint32_t mx = (uint32_t)std::max(imgWidth, imgHeight);
    const uint32_t cw = div_up(mx, cellCols);
    const uint32_t c = div_up((uint32_t)imgWidth, cw), r = div_up((uint32_t)imgHeight, cw);
    const uint32_t nCells = c * r;
    RSFM_RETURN_IF(nbMatches > 0 && (matches == nullptr || cols == nullptr || rows == nullptr || cellWidth == nullptr ||
                                     reordered == nullptr || bounds == nullptr || order == nullptr),
                   RSFM_ERR_INVALID, "rmf bin: null argument");
    for (uint32_t k = 0; k < nbMatches; k++)
        RSFM_RETURN_IF(!std::isfinite(matches[k].x0) || !std::isfinite(matches[k].y0), RSFM_ERR_INVALID,
                       "rmf bin: non-finite keypoint coordinate at match " + std::to_string(k));
    std::vector<uint32_t> cell(nbMatches), cnt(nCells + 1, 0);
    for (uint32_t k = 0; k < nbMatches; k++) {
        const float fi = std::min(std::max(std::ceil(matches[k].y0 / (float)cw), 0.f), (float)r - 1.f);
        const float fj = std::min(std::max(std::ceil(matches[k].x0 / (float)cw), 0.f), (float)c - 1.f);
        cell[k] = (uint32_t)fi * c + (uint32_t)fj;
        cnt[cell[k] + 1]++;
    }
    for (uint32_t i = 0; i < nCells; i++) cnt[i + 1] += cnt[i];
    std::copy(cnt.begin(), cnt.end(), bounds);
    for (uint32_t k = 0; k < nbMatches; k++) {
        const uint32_t p = cnt[cell[k]]++;
        reordered[p] = matches[k];
        order[p] = k;
    }
    *cols = c;
    *rows = r;
    *cellWidth = cw;
    return RSFM_OK;
}
