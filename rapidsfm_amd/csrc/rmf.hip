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
#include "internal.hpp"
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

// one cell (bx, by): the best of nbTests affine hypotheses
__device__ inline void best_affine_cell(Reducer& R, float* __restrict__ affine, const float4* __restrict__ m,
                                        const uint32_t* __restrict__ bounds, uint32_t cols, uint32_t rows,
                                        uint32_t nbTests, float threshold, uint32_t seed, uint32_t bx, uint32_t by)
{
    Cells c;
    cells_init(c, bounds, cols, rows, bx, by);
    const uint32_t rank = by * cols + bx;
    if (c.sum < 9) {
        if (threadIdx.x < 6) affine[rank * 6 + threadIdx.x] = __builtin_nanf("");
        return;
    }
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

__global__ __launch_bounds__(kCta) void best_affine_kernel(float* __restrict__ affine, const float4* __restrict__ m,
                                                           const uint32_t* __restrict__ bounds, uint32_t cols,
                                                           uint32_t rows, uint32_t nbTests, float threshold, uint32_t seed)
{
    __shared__ Reducer R;
    best_affine_cell(R, affine, m, bounds, cols, rows, nbTests, threshold, seed, blockIdx.x, blockIdx.y);
}

__device__ inline void vote(uint32_t* votesWords, uint32_t idx)
{
    atomicAdd(votesWords + (idx >> 2), 1u << (8 * (idx & 3)));
}

struct VoteSmem {
    Reducer R;
    float best[6];
    uint32_t bestCnt;
};

// one cell: its own affine (or, with tryOther, the best of every cell's affine
// on its neighbourhood) votes for its inliers when it explains >= 6 of them
__device__ inline void vote_cell(VoteSmem& S, uint32_t* __restrict__ votesWords, const float* __restrict__ affine,
                                 const float4* __restrict__ m, const uint32_t* __restrict__ bounds, uint32_t cols,
                                 uint32_t rows, float threshold, bool tryOther, uint32_t bx, uint32_t by)
{
    Cells c;
    cells_init(c, bounds, cols, rows, bx, by);
    const uint32_t rank = by * cols + bx;
    const uint32_t nCells = cols * rows;
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
        cta_find_max(S.R, bestCnt, best);
    } else if (threadIdx.x == 0) {
        bestCnt = count_inliers(c, best, m, threshold);
    }
    if (threadIdx.x == 0) {
        S.bestCnt = bestCnt;
#pragma unroll
        for (int q = 0; q < 6; q++) S.best[q] = best[q];
    }
    __syncthreads();
    if (S.bestCnt < kMinInliers) return;
    float T[6];
#pragma unroll
    for (int q = 0; q < 6; q++) T[q] = S.best[q];
    const float sq = sq_threshold(T, threshold);
    for (uint32_t k = threadIdx.x; k < c.sum; k += kCta) {
        const uint32_t mi = cells_map(c, k);
        if (is_inlier(T, m[mi], sq)) vote(votesWords, mi);
    }
}

__global__ __launch_bounds__(kCta) void vote_kernel(uint32_t* __restrict__ votesWords, const float* __restrict__ affine,
                                                    const float4* __restrict__ m, const uint32_t* __restrict__ bounds,
                                                    uint32_t cols, uint32_t rows, float threshold, bool tryOther)
{
    __shared__ VoteSmem S;
    vote_cell(S, votesWords, affine, m, bounds, cols, rows, threshold, tryOther, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Batched form (no reference counterpart: the reference bins each pair's
// matches on the host and runs one pair per call with a blocking sync,
// RansacMatchFilter.cpp:42-133). Per pair: device binning with the
// reference's rule and order, the two cell kernels over all pairs in one
// launch each, then the mask (votes >= minVotes) and an in-place stable
// compaction of the gathered pairs and matches with the kept count.
constexpr int kBinThreads = 1024;
constexpr int kBinWaves = kBinThreads / 64;
constexpr uint32_t kMaxCells = 1024;

__device__ inline uint32_t task_n(const BatchTask& T) { return min(*T.count, T.nHost); }

// cell of a point: the reference's ceil rule, clamped (RansacMatchFilter.cpp:47-52)
__device__ inline uint32_t cell_of(float x, float y, const BatchTask& T)
{
    const float cw = (float)T.cellWidth;
    const float fi = fminf(fmaxf(ceilf(y / cw), 0.f), (float)T.rows - 1.f);
    const float fj = fminf(fmaxf(ceilf(x / cw), 0.f), (float)T.cols - 1.f);
    return (uint32_t)fi * T.cols + (uint32_t)fj;
}

struct BinSmem {
    uint32_t cnt[kMaxCells];
    uint32_t cursor[kMaxCells];
    uint16_t waveCnt[kBinWaves][kMaxCells];
    uint32_t scan[kBinThreads];
};

// one workgroup per task: votes cleared, cell counts, bounds (exclusive
// prefix, cells in row-major order), then a stable scatter: points in index
// order, a point's slot = its cell's start + earlier points of the same cell
// (ballot peer groups inside a wave, per-wave cell counts across the waves of a
// round of 1024 points, a running cursor across rounds)
template <class Src>
__global__ __launch_bounds__(kBinThreads) void bin_kernel(const Src batch)
{
    __shared__ BinSmem S;
    const BatchTask& T = batch[blockIdx.x];
    const uint32_t n = task_n(T);
    const uint32_t nCells = T.cols * T.rows;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t w = tid; w < (T.nHost + 3) / 4; w += kBinThreads) reinterpret_cast<uint32_t*>(T.votes)[w] = 0;
    for (uint32_t c = tid; c < kMaxCells; c += kBinThreads) S.cnt[c] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kBinThreads) {
        const float4 p = T.pts[i];
        atomicAdd(&S.cnt[cell_of(p.x, p.y, T)], 1u);
    }
    __syncthreads();
    // exclusive prefix over the cells (nCells <= kMaxCells = kBinThreads)
    const uint32_t v = tid < nCells ? S.cnt[tid] : 0u;
    S.scan[tid] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kBinThreads; off <<= 1) {
        const uint32_t a = tid >= off ? S.scan[tid - off] : 0u;
        __syncthreads();
        S.scan[tid] += a;
        __syncthreads();
    }
    if (tid < nCells) {
        const uint32_t start = S.scan[tid] - v;
        S.cursor[tid] = start;
        T.bounds[tid] = start;
    }
    if (tid == 0) T.bounds[nCells] = n;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kBinThreads) {
        const uint32_t i = base + tid;
        const bool valid = i < n;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t cell = ~0u;
        if (valid) {
            p = T.pts[i];
            cell = cell_of(p.x, p.y, T);
        }
        for (uint32_t c = lane; c < nCells; c += 64) S.waveCnt[wave][c] = 0;
        // lanes of this wave with the same cell
        uint64_t peers = 0;
        uint64_t todo = __ballot(valid);
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const uint32_t lc = (uint32_t)__shfl((int)cell, leader);
            const uint64_t same = __ballot(valid && cell == lc);
            if (valid && cell == lc) peers = same;
            todo &= ~same;
        }
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        __syncthreads();  // waveCnt cleared by every wave
        if (valid && rank == 0) S.waveCnt[wave][cell] = (uint16_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = S.cursor[cell] + rank;
            for (uint32_t w = 0; w < wave; w++) pos += S.waveCnt[w][cell];
            T.reordered[pos] = p;
            T.pos[i] = pos;
        }
        __syncthreads();
        for (uint32_t c = tid; c < nCells; c += kBinThreads) {
            uint32_t add = 0;
            for (uint32_t w = 0; w < kBinWaves; w++) add += S.waveCnt[w][c];
            S.cursor[c] += add;
        }
        __syncthreads();
    }
}

template <class Src>
__global__ __launch_bounds__(kCta) void best_affine_batch_kernel(const Src batch, BatchParams prm)
{
    __shared__ Reducer R;
    const BatchTask& T = batch[blockIdx.z];
    if (blockIdx.x >= T.cols || blockIdx.y >= T.rows || task_n(T) < 16) return;
    best_affine_cell(R, T.affine, T.reordered, T.bounds, T.cols, T.rows, prm.nbTests, T.threshold, T.seed, blockIdx.x,
                     blockIdx.y);
}

template <class Src>
__global__ __launch_bounds__(kCta) void vote_batch_kernel(const Src batch, BatchParams prm)
{
    __shared__ VoteSmem S;
    const BatchTask& T = batch[blockIdx.z];
    if (blockIdx.x >= T.cols || blockIdx.y >= T.rows || task_n(T) < 16) return;
    vote_cell(S, reinterpret_cast<uint32_t*>(T.votes), T.affine, T.reordered, T.bounds, T.cols, T.rows, T.threshold,
              prm.tryOther != 0, blockIdx.x, blockIdx.y);
}

// one workgroup per task: keep[i] = votes of point i >= minVotes; the kept
// points (and matches) move to the front in index order, in place (a round
// reads all of its 1024 entries before any is written, and writes only below
// the next round's first entry); the count becomes the kept count
template <class Src>
__global__ __launch_bounds__(kBinThreads) void compact_kernel(const Src batch, BatchParams prm)
{
    __shared__ uint32_t sWave[kBinWaves];
    __shared__ uint32_t sBase;
    const BatchTask& T = batch[blockIdx.x];
    const uint32_t n = task_n(T);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) sBase = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += kBinThreads) {
        const uint32_t i = base + tid;
        bool keep = false;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        rsfm_valid_match mv{};
        if (i < n) {
            keep = n >= 16 && T.votes[T.pos[i]] >= prm.minVotes;
            p = T.pts[i];
            if (T.matches) mv = T.matches[i];
            if (T.mask) T.mask[i] = keep ? 1 : 0;
        }
        const uint64_t b = __ballot(keep);
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (lane == 0) sWave[wave] = (uint32_t)__popcll(b);
        __syncthreads();
        uint32_t off = sBase;
        for (uint32_t w = 0; w < wave; w++) off += sWave[w];
        if (keep) {
            const uint32_t o = off + (uint32_t)__popcll(b & below);
            T.pts[o] = p;
            if (T.matches) T.matches[o] = mv;
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < kBinWaves; w++) t += sWave[w];
            sBase += t;
        }
        __syncthreads();
    }
    if (T.mask)
        for (uint32_t i = n + tid; i < T.nHost; i += kBinThreads) T.mask[i] = 0;
    if (tid == 0) *T.count = sBase;
}

int launch_batch_table(const BatchTask* tasks, uint32_t nTasks, uint32_t maxCols, uint32_t maxRows, const BatchParams& p,
                       hipStream_t stream)
{
    if (nTasks == 0) return RSFM_OK;
    const InTable<BatchTask> src{tasks};
    bin_kernel<<<nTasks, kBinThreads, 0, stream>>>(src);
    RSFM_CHECK(check_launch("rmf bin table"));
    best_affine_batch_kernel<<<dim3(maxCols, maxRows, nTasks), kCta, 0, stream>>>(src, p);
    RSFM_CHECK(check_launch("rmf best affine table"));
    vote_batch_kernel<<<dim3(maxCols, maxRows, nTasks), kCta, 0, stream>>>(src, p);
    RSFM_CHECK(check_launch("rmf vote table"));
    compact_kernel<<<nTasks, kBinThreads, 0, stream>>>(src, p);
    return check_launch("rmf compact table");
}

size_t batch_workspace_bytes(uint32_t nPts, uint32_t nCells)
{
    return align_up((size_t)nPts * 16, 256) + align_up((size_t)nPts * 4, 256) + align_up(((size_t)nCells + 1) * 4, 256) +
           align_up((size_t)nCells * 24, 256) + align_up(((size_t)nPts + 3) / 4 * 4, 256);
}

void place_workspace(BatchTask& t, uint8_t*& ws)
{
    const uint32_t nCells = t.cols * t.rows;
    t.reordered = reinterpret_cast<float4*>(ws);
    ws += align_up((size_t)t.nHost * 16, 256);
    t.pos = reinterpret_cast<uint32_t*>(ws);
    ws += align_up((size_t)t.nHost * 4, 256);
    t.bounds = reinterpret_cast<uint32_t*>(ws);
    ws += align_up(((size_t)nCells + 1) * 4, 256);
    t.affine = reinterpret_cast<float*>(ws);
    ws += align_up((size_t)nCells * 24, 256);
    t.votes = ws;
    ws += align_up(((size_t)t.nHost + 3) / 4 * 4, 256);
}

int make_batch_task(const rsfm_rmf_task& t, uint32_t cellCols, float relativeThreshold, BatchTask& out)
{
    RSFM_RETURN_IF(t.imgWidth <= 0 || t.imgHeight <= 0, RSFM_ERR_INVALID, "rmf batch: bad image size");
    RSFM_RETURN_IF(t.nbPtsDev == nullptr, RSFM_ERR_INVALID, "rmf batch: nbPtsDev (device count) is required");
    RSFM_RETURN_IF(t.nbPts > 0 && t.pts == nullptr, RSFM_ERR_INVALID, "rmf batch: pts is null");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16, RSFM_ERR_INVALID, "rmf batch: pts must be 16-byte aligned");
    const uint32_t mx = (uint32_t)std::max(t.imgWidth, t.imgHeight);
    const uint32_t cw = div_up(mx, cellCols);
    const uint32_t cols = div_up((uint32_t)t.imgWidth, cw), rows = div_up((uint32_t)t.imgHeight, cw);
    RSFM_RETURN_IF(cols * rows > kMaxCells, RSFM_ERR_INVALID, "rmf batch: more than 1024 cells");
    out = BatchTask{};
    out.pts = reinterpret_cast<float4*>(t.pts);
    out.matches = t.matches;
    out.count = t.nbPtsDev;
    out.nHost = t.nbPts;
    out.mask = t.inlierMask;
    out.cols = cols;
    out.rows = rows;
    out.cellWidth = cw;
    out.threshold = (float)cw * relativeThreshold;
    out.seed = t.seed;
    return RSFM_OK;
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

extern "C" int rsfm_rmf_filter_batch_workspace_size(const rsfm_rmf_task* tasks, size_t nbTasks, uint32_t cellCols,
                                                    size_t* bytes)
{
    RSFM_RETURN_IF(bytes == nullptr || (nbTasks > 0 && tasks == nullptr), RSFM_ERR_INVALID, "rmf batch: null argument");
    RSFM_RETURN_IF(cellCols == 0, RSFM_ERR_INVALID, "rmf batch: cellCols must be >= 1");
    size_t total = 0;
    for (size_t i = 0; i < nbTasks; i++) {
        rmf::BatchTask t;
        RSFM_CHECK(rmf::make_batch_task(tasks[i], cellCols, 1.f, t));
        total += rmf::batch_workspace_bytes(t.nHost, t.cols * t.rows);
    }
    *bytes = total;
    return RSFM_OK;
}

extern "C" int rsfm_rmf_filter_batch(const rsfm_rmf_task* tasks, size_t nbTasks, uint32_t cellCols,
                                     uint32_t nbRansacTests, float relativeThreshold, uint32_t minVotes,
                                     int tryOtherAffines, void* workspace, size_t workspaceBytes, hipStream_t stream)
{
    using namespace rmf;
    size_t need = 0;
    RSFM_CHECK(rsfm_rmf_filter_batch_workspace_size(tasks, nbTasks, cellCols, &need));
    RSFM_RETURN_IF(workspaceBytes < need, RSFM_ERR_INVALID,
                   "rmf batch: workspace too small (" + std::to_string(workspaceBytes) + " < " + std::to_string(need) + ")");
    RSFM_RETURN_IF(need > 0 && (workspace == nullptr || reinterpret_cast<uintptr_t>(workspace) % 256),
                   RSFM_ERR_INVALID, "rmf batch: workspace must be non-null and 256-byte aligned");
    RSFM_RETURN_IF(!(relativeThreshold > 0.f), RSFM_ERR_INVALID, "rmf batch: threshold must be > 0");
    const BatchParams prm{nbRansacTests, minVotes, tryOtherAffines};
    uint8_t* ws = static_cast<uint8_t*>(workspace);
    constexpr size_t kMax = 32;
    for (size_t b = 0; b < nbTasks; b += kMax) {
        const size_t n = std::min(kMax, nbTasks - b);
        ByValue<BatchTask, kMax> batch{};
        uint32_t maxCols = 0, maxRows = 0;
        for (size_t i = 0; i < n; i++) {
            RSFM_CHECK(make_batch_task(tasks[b + i], cellCols, relativeThreshold, batch.t[i]));
            place_workspace(batch.t[i], ws);
            maxCols = std::max(maxCols, batch.t[i].cols);
            maxRows = std::max(maxRows, batch.t[i].rows);
        }
        bin_kernel<<<(uint32_t)n, kBinThreads, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("rmf bin"));
        best_affine_batch_kernel<<<dim3(maxCols, maxRows, (uint32_t)n), kCta, 0, stream>>>(batch, prm);
        RSFM_CHECK(check_launch("rmf best affine batch"));
        vote_batch_kernel<<<dim3(maxCols, maxRows, (uint32_t)n), kCta, 0, stream>>>(batch, prm);
        RSFM_CHECK(check_launch("rmf vote batch"));
        compact_kernel<<<(uint32_t)n, kBinThreads, 0, stream>>>(batch, prm);
        RSFM_CHECK(check_launch("rmf compact"));
    }
    return RSFM_OK;
}
