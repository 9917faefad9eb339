// ransac.hip - batched fundamental-matrix RANSAC hypothesis fitting and scoring.
//
// Reference path: ransac<F,8> (ransac.hpp:115-163) driving
// findEpipolarityRansacWithInsufficientTiePts<0> (ImagePair.cpp:148-204):
// per hypothesis the sample sanity gate (ImagePair.cpp:160-177), the 8-point
// fit (legacy/geometry.hpp:216-259) and the inlier count of
// checkEpipolarTransformation (legacy/geometry.cpp:55-86) over all points.
// Sample indices are an input (the reference draws them from a
// random_device-seeded engine), so a fixed index set gives one defined
// inlier count per hypothesis.
//
// GPU layout: a 256-thread workgroup owns 256 hypotheses of one pair. Each
// lane fits one hypothesis in registers (fgeom::fit8) into LDS; then each
// wavefront scores its 64 hypotheses one at a time, lanes striding over the
// correspondences staged in LDS tiles of 2048, counting inliers with
// ballot + popcount on the scalar unit.
#include "internal.hpp"
#include "fgeom.hpp"
#include "hgeom.hpp"
#include <algorithm>

namespace rsfm_amd {
namespace ransac {

constexpr int kThreads = 256;
constexpr int kMaxTasks = 32;

typedef ByValue<DevTask, kMaxTasks> Batch;

__device__ inline uint32_t task_n(const DevTask& T) { return T.nDev ? min(*T.nDev, T.nHost) : T.nHost; }

// Model kinds of the hypothesis kernels: 0 / 1 = fundamental matrix from 8
// points with the two-sided / Sampson test, 2 = homography from 4 points with
// checkProjectiveTransformation's test.
template <int kMode>
constexpr int sample_size() { return kMode == 2 ? 4 : 8; }

template <int kMode>
__device__ inline bool model_inlier(const float M[9], float4 p, float t2)
{
    if constexpr (kMode == 2) return hgeom::inlier_h(M, p, t2);
    else return fgeom::inlier<kMode>(M, p, t2);
}

// sanity gate of ImagePair.cpp:167-177 (and :366-374 for H): any two samples
// closer than the threshold in either image reject the hypothesis
template <int kN>
__device__ inline bool sane(const float4 p[kN], float s0, float s1)
{
    if (s0 <= 0.f && s1 <= 0.f) return true;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < kN; i++)
#pragma unroll
        for (int j = 0; j < i; j++) {
            const float dx0 = p[i].x - p[j].x, dy0 = p[i].y - p[j].y;
            const float dx1 = p[i].z - p[j].z, dy1 = p[i].w - p[j].w;
            if (dx0 * dx0 + dy0 * dy0 < s0 || dx1 * dx1 + dy1 * dy1 < s1) ok = false;
        }
    return ok;
}

// fit of hypothesis h; false when out of range, insane or degenerate
template <int kMode>
__device__ inline bool fit_hypothesis(const DevTask& T, uint32_t n, uint32_t h, float s0, float s1, float M[9])
{
    constexpr int kN = sample_size<kMode>();
    if (h >= T.H || n < (uint32_t)kN) return false;
    float4 p[kN];
    bool inRange = true;
#pragma unroll
    for (int i = 0; i < kN; i++) {
        const uint32_t k = T.idx[(size_t)h * kN + i];
        inRange = inRange && k < n;
        p[i] = T.pts[k < n ? k : 0];
    }
    if (!inRange || !sane<kN>(p, s0, s1)) return false;
    if constexpr (kMode == 2) return fgeom::fit_h4(p, M);
    else return fgeom::fit8(p, M);
}

// Scoring: one lane per hypothesis; a workgroup holds 64 hypotheses and
// kScoreWaves wavefronts. Wavefront 0 fits the 64 models in registers and
// shares them through LDS; then each wavefront walks its own contiguous part
// of the correspondences in index order (all lanes read the same
// correspondence, so the loop body is VALU work only), and the partial counts
// are summed through LDS. Splitting the points keeps enough wavefronts per
// SIMD to hide load and dependency latency: one wavefront per 64 hypotheses
// gives only two per SIMD for 32 pairs of 4096 hypotheses.
constexpr int kScoreThreads = 64;
#ifndef RSFM_SCORE_WAVES
#define RSFM_SCORE_WAVES 4
#endif
constexpr int kScoreWaves = RSFM_SCORE_WAVES;
constexpr int kScoreBlock = kScoreThreads * kScoreWaves;

// points [beg, end) of wavefront w: contiguous parts starting on multiples of 8
__device__ inline void wave_range(uint32_t n, uint32_t w, uint32_t& beg, uint32_t& end)
{
    const uint32_t part = ((n + kScoreWaves - 1) / kScoreWaves + 7u) & ~7u;
    beg = min(n, w * part);
    end = min(n, beg + part);
}

// number of inliers of model F among P[0, m): the next eight points are in
// flight while the current eight are tested
template <int kMode>
__device__ inline uint32_t count_inliers(const float F[9], const float4* __restrict__ P, uint32_t m, float t2)
{
    auto test4 = [&](const float4& a, const float4& b, const float4& c, const float4& d) -> uint32_t {
        return (model_inlier<kMode>(F, a, t2) ? 1u : 0u) + (model_inlier<kMode>(F, b, t2) ? 1u : 0u) +
               (model_inlier<kMode>(F, c, t2) ? 1u : 0u) + (model_inlier<kMode>(F, d, t2) ? 1u : 0u);
    };
    uint32_t cnt = 0, i = 0;
    if (m >= 8) {
        float4 c0 = P[0], c1 = P[1], c2 = P[2], c3 = P[3], c4 = P[4], c5 = P[5], c6 = P[6], c7 = P[7];
        for (; i + 16 <= m; i += 8) {
            const float4* __restrict__ Q = P + i + 8;  // one address per trip, immediate offsets
            const float4 n0 = Q[0], n1 = Q[1], n2 = Q[2], n3 = Q[3];
            const float4 n4 = Q[4], n5 = Q[5], n6 = Q[6], n7 = Q[7];
            cnt += test4(c0, c1, c2, c3) + test4(c4, c5, c6, c7);
            c0 = n0;
            c1 = n1;
            c2 = n2;
            c3 = n3;
            c4 = n4;
            c5 = n5;
            c6 = n6;
            c7 = n7;
        }
        cnt += test4(c0, c1, c2, c3) + test4(c4, c5, c6, c7);
        i += 8;
    }
    for (; i < m; i++) cnt += model_inlier<kMode>(F, P[i], t2) ? 1u : 0u;
    return cnt;
}

template <int kMode, class Src>
__global__ __launch_bounds__(kScoreBlock) void score_kernel(const Src batch, float t2, float s0, float s1)
{
    const Wg3 wg = xcd_local_block();  // all hypothesis blocks of a pair share one L2
    const DevTask& T = batch[wg.y];
    if (wg.x * kScoreThreads >= T.H) return;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t h = wg.x * kScoreThreads + lane;
    const uint32_t n = task_n(T);
    __shared__ float sF[9][kScoreThreads];
    __shared__ uint32_t sOk[kScoreThreads];
    __shared__ uint32_t sCnt[kScoreWaves][kScoreThreads];
    float F[9];
    if (w == 0) {
        const bool ok = fit_hypothesis<kMode>(T, n, h, s0, s1, F);
#pragma unroll
        for (int q = 0; q < 9; q++) sF[q][lane] = ok ? F[q] : 0.f;
        sOk[lane] = ok ? 1u : 0u;
    }
    __syncthreads();
    const bool ok = sOk[lane] != 0;
#pragma unroll
    for (int q = 0; q < 9; q++) F[q] = sF[q][lane];
    uint32_t cnt = 0;
    if (ok) {
        uint32_t beg, end;
        wave_range(n, w, beg, end);
        // the offset is hidden from the uniformity analysis so the loads take
        // the vector memory path (in-order vmcnt, so the prefetch overlaps);
        // all 64 lanes read the same address, one cache line per wavefront
        uint32_t base = beg;
        asm volatile("" : "+v"(base));
        cnt = count_inliers<kMode>(F, T.pts + base, end - beg, t2);
    }
    sCnt[w][lane] = cnt;
    __syncthreads();
    if (w != 0 || h >= T.H) return;
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < kScoreWaves; k++) total += sCnt[k][lane];
    T.counts[h] = ok ? total : 0u;
    if (T.models) {
#pragma unroll
        for (int q = 0; q < 9; q++) T.models[(size_t)h * 9 + q] = F[q];
    }
}

// best hypothesis (largest count, lowest index among equals) and its inlier
// mask: 16 wavefronts per task, the maximum key through wave shuffles and one
// LDS round, the mask four points per thread and trip with the loads first
constexpr int kBestThreads = 1024;

template <int kMode, class Src>
__global__ __launch_bounds__(kBestThreads) void best_kernel(const Src batch, float t2, float s0, float s1)
{
    const DevTask& T = batch[blockIdx.x];
    const uint32_t n = task_n(T);
    constexpr int kWaves = kBestThreads / 64;
    __shared__ uint64_t sKey[kWaves];
    __shared__ float sF[9];
    __shared__ uint32_t sOk;
    const uint32_t tid = threadIdx.x;
    uint64_t key = 0;  // (count << 32) | ~h : max count, then lowest h
    for (uint32_t h = tid; h < T.H; h += kBestThreads) {
        const uint64_t k = ((uint64_t)T.counts[h] << 32) | (uint32_t)~h;
        key = k > key ? k : key;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t k = __shfl_xor(key, o);
        key = k > key ? k : key;
    }
    if ((tid & 63) == 0) sKey[tid >> 6] = key;
    __syncthreads();
    if (tid == 0) {
        uint64_t bk = 0;
        for (int w = 0; w < kWaves; w++) bk = sKey[w] > bk ? sKey[w] : bk;
        const uint32_t bestH = T.H ? ~(uint32_t)bk : 0u;
        const uint32_t bestC = (uint32_t)(bk >> 32);
        float F[9];
        bool ok = T.H > 0 && bestC > 0;
        if (T.models != nullptr) {
            // the scorer stored every model; a nonzero count means its fit succeeded
            if (ok)
                for (int q = 0; q < 9; q++) F[q] = T.models[(size_t)bestH * 9 + q];
        } else {
            ok = ok && fit_hypothesis<kMode>(T, n, bestH, s0, s1, F);
        }
        for (int q = 0; q < 9; q++) sF[q] = ok ? F[q] : 0.f;
        sOk = ok;
        if (T.best) {
            T.best[0] = T.H ? bestH : ~0u;
            T.best[1] = bestC;
        }
    }
    __syncthreads();
    if (!T.mask) return;
    float F[9];
    for (int q = 0; q < 9; q++) F[q] = sF[q];
    const bool ok = sOk != 0;
    for (uint32_t i0 = tid; i0 < T.nHost; i0 += 4 * kBestThreads) {
        float4 p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = i0 + u * kBestThreads;
            if (ok && i < n) p[u] = T.pts[i];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = i0 + u * kBestThreads;
            if (i < T.nHost) T.mask[i] = (ok && i < n) ? (uint8_t)model_inlier<kMode>(F, p[u], t2) : 0;
        }
    }
}

// ---------------------------------------------------------------------------
// 7-point hypotheses (the sampling model of findEpipolarityDegensac,
// degensac.cpp:113-150): each sample of 7 indices gives up to three F
// (fgeom::fit7); a hypothesis scores as its best model (most inliers, first
// on ties), which is also the model reported for it.
constexpr int kMaxF7 = 3;

__device__ inline bool sane7(const float4 p[7], float s0, float s1)
{
    if (s0 <= 0.f && s1 <= 0.f) return true;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 7; i++)
#pragma unroll
        for (int j = 0; j < i; j++) {
            const float dx0 = p[i].x - p[j].x, dy0 = p[i].y - p[j].y;
            const float dx1 = p[i].z - p[j].z, dy1 = p[i].w - p[j].w;
            if (dx0 * dx0 + dy0 * dy0 < s0 || dx1 * dx1 + dy1 * dy1 < s1) ok = false;
        }
    return ok;
}

__device__ inline int fit_hypothesis7(const DevTask& T, uint32_t n, uint32_t h, float s0, float s1, float Fs[kMaxF7][9])
{
    if (h >= T.H || n < 7) return 0;
    float4 p[7];
    bool inRange = true;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const uint32_t k = T.idx[(size_t)h * 7 + i];
        inRange = inRange && k < n;
        p[i] = T.pts[k < n ? k : 0];
    }
    if (!inRange || !sane7(p, s0, s1)) return 0;
    return fgeom::fit7(p, Fs);
}

// 7-point scoring. A workgroup takes kS7Samples samples, one per lane of its
// first two wavefronts: every such lane solves its sample (up to three F) and
// the workgroup's M models are listed in LDS in sample order (a scan of the
// model counts). Lane l of every wavefront then holds models l, l + 64, ...
// (R = ceil(M / 64) of them, at most six) in registers and tests each point
// of its quarter of the correspondences against all R, as score_kernel walks
// them; the four partial counts are added with LDS atomics. A sample with one
// real root thus takes one model slot instead of three (about 2.5 models per
// sample with 8K-keypoint pairs: R = 5 instead of six slots per lane), and
// every point load feeds R tests. Each sample then reports its best model
// (most inliers, the first on ties) and that count.
constexpr uint32_t kS7Samples = 128;
constexpr int kS7MaxRounds = (int)(kS7Samples * kMaxF7 / 64);
constexpr uint32_t kS7MaxModels = kS7Samples * kMaxF7;
static_assert(kS7Samples <= kScoreBlock && kS7Samples % 64 == 0, "samples are fitted one per lane");

// inliers of the lane's R models among P[0, m): the next four points are in
// flight while the current four are tested (four, not eight: the R models
// already hold up to 54 registers, and the kernel keeps four waves per SIMD)
template <int kMode, int R>
__device__ inline void count7(const float (&F)[kS7MaxRounds][9], const float4* __restrict__ P, uint32_t m, float t2,
                              uint32_t (&c)[kS7MaxRounds])
{
    auto test = [&](const float4& p) {
#pragma unroll
        for (int r = 0; r < R; r++) c[r] += fgeom::inlier<kMode>(F[r], p, t2) ? 1u : 0u;
    };
    uint32_t i = 0;
    if (m >= 4) {
        float4 q0 = P[0], q1 = P[1], q2 = P[2], q3 = P[3];
        for (; i + 8 <= m; i += 4) {
            const float4* __restrict__ Q = P + i + 4;
            const float4 n0 = Q[0], n1 = Q[1], n2 = Q[2], n3 = Q[3];
            test(q0);
            test(q1);
            test(q2);
            test(q3);
            q0 = n0;
            q1 = n1;
            q2 = n2;
            q3 = n3;
        }
        test(q0);
        test(q1);
        test(q2);
        test(q3);
        i += 4;
    }
    for (; i < m; i++) test(P[i]);
}

template <int kMode, class Src>
__global__ __launch_bounds__(kScoreBlock) void score7_kernel(const Src batch, float t2, float s0, float s1)
{
    const Wg3 wg = xcd_local_block();
    const DevTask& T = batch[wg.y];
    if (wg.x * kS7Samples >= T.H) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t h = wg.x * kS7Samples + tid;
    const uint32_t n = task_n(T);
    __shared__ float sF[9][kS7MaxModels];
    __shared__ uint32_t sCnt[kS7MaxModels];
    __shared__ uint32_t sWaveModels[kS7Samples / 64];
    for (uint32_t k = tid; k < kS7MaxModels; k += kScoreBlock) sCnt[k] = 0u;
    float Fw[kMaxF7][9];  // written by the solver with a running index
    const int nm = tid < kS7Samples ? fit_hypothesis7(T, n, h, s0, s1, Fw) : 0;
    // the models' place in the list: prefix of nm in lane order (the lanes
    // past the samples add nothing)
    uint32_t inc = (uint32_t)nm;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)inc, o);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63 && w < kS7Samples / 64) sWaveModels[w] = inc;
    __syncthreads();
    uint32_t off = inc - (uint32_t)nm;
#pragma unroll
    for (uint32_t k = 0; k < kS7Samples / 64; k++) off += k < w ? sWaveModels[k] : 0u;
#pragma unroll
    for (int mi = 0; mi < kMaxF7; mi++)
        if (mi < nm)
#pragma unroll
            for (int q = 0; q < 9; q++) sF[q][off + mi] = Fw[mi][q];
    __syncthreads();
    uint32_t M = 0;
#pragma unroll
    for (uint32_t k = 0; k < kS7Samples / 64; k++) M += sWaveModels[k];
    const int R = (int)((M + 63) / 64);
    float F[kS7MaxRounds][9];  // model r * 64 + lane; zero (never an inlier) past the list
#pragma unroll
    for (int r = 0; r < kS7MaxRounds; r++) {
        const uint32_t k = (uint32_t)r * 64 + lane;
#pragma unroll
        for (int q = 0; q < 9; q++) F[r][q] = k < M ? sF[q][k] : 0.f;
    }
    uint32_t c[kS7MaxRounds] = {0u, 0u, 0u, 0u, 0u, 0u};
    uint32_t beg, end;
    wave_range(n, w, beg, end);
    // the offset is hidden from the uniformity analysis so the loads take the
    // vector memory path (see score_kernel)
    uint32_t base = beg;
    asm volatile("" : "+v"(base));
    const float4* P = T.pts + base;
    const uint32_t m = end - beg;
    switch (R) {
        case 1: count7<kMode, 1>(F, P, m, t2, c); break;
        case 2: count7<kMode, 2>(F, P, m, t2, c); break;
        case 3: count7<kMode, 3>(F, P, m, t2, c); break;
        case 4: count7<kMode, 4>(F, P, m, t2, c); break;
        case 5: count7<kMode, 5>(F, P, m, t2, c); break;
        case 6: count7<kMode, 6>(F, P, m, t2, c); break;
        default: break;
    }
#pragma unroll
    for (int r = 0; r < kS7MaxRounds; r++) {
        const uint32_t k = (uint32_t)r * 64 + lane;
        if (r < R && k < M) atomicAdd(&sCnt[k], c[r]);
    }
    __syncthreads();
    if (tid >= kS7Samples || h >= T.H) return;
    uint32_t best = 0;
    int bm = 0;
    for (int mi = 0; mi < nm; mi++) {
        const uint32_t cm = sCnt[off + mi];
        if (mi == 0 || cm > best) {
            best = cm;
            bm = mi;
        }
    }
    T.counts[h] = nm ? best : 0u;
    if (T.models) {
#pragma unroll
        for (int q = 0; q < 9; q++) T.models[(size_t)h * 9 + q] = nm ? sF[q][off + bm] : 0.f;
    }
}

// best 7-point hypothesis (largest count, lowest index), its best model and mask
template <int kMode, class Src>
__global__ __launch_bounds__(kThreads) void best7_kernel(const Src batch, float t2, float s0, float s1)
{
    const DevTask& T = batch[blockIdx.x];
    const uint32_t n = task_n(T);
    __shared__ uint64_t sKey[kThreads];
    __shared__ uint32_t sRed[kMaxF7][kThreads];
    __shared__ float sFs[kMaxF7][9];
    __shared__ int sNm;
    uint64_t key = 0;
    for (uint32_t h = threadIdx.x; h < T.H; h += kThreads) {
        const uint64_t k = ((uint64_t)T.counts[h] << 32) | (uint32_t)~h;
        key = k > key ? k : key;
    }
    sKey[threadIdx.x] = key;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sKey[threadIdx.x] = max(sKey[threadIdx.x], sKey[threadIdx.x + s]);
        __syncthreads();
    }
    const uint64_t bk = sKey[0];
    const uint32_t bestH = T.H ? ~(uint32_t)bk : 0u;
    const uint32_t bestC = (uint32_t)(bk >> 32);
    if (threadIdx.x == 0) {
        float Fs[kMaxF7][9];
        const int nm = (T.H > 0 && bestC > 0) ? fit_hypothesis7(T, n, bestH, s0, s1, Fs) : 0;
        for (int m = 0; m < kMaxF7; m++)
            for (int q = 0; q < 9; q++) sFs[m][q] = m < nm ? Fs[m][q] : 0.f;
        sNm = nm;
    }
    __syncthreads();
    const int nm = sNm;
    // counts of the (up to three) models of the best sample
    uint32_t c[kMaxF7] = {0, 0, 0};
    for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
        const float4 p = T.pts[i];
        for (int m = 0; m < nm; m++) c[m] += fgeom::inlier<kMode>(sFs[m], p, t2) ? 1u : 0u;
    }
    for (int m = 0; m < kMaxF7; m++) sRed[m][threadIdx.x] = c[m];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int m = 0; m < kMaxF7; m++) sRed[m][threadIdx.x] += sRed[m][threadIdx.x + s];
        __syncthreads();
    }
    int bm = 0;
    for (int m = 1; m < nm; m++)
        if (sRed[m][0] > sRed[bm][0]) bm = m;
    if (threadIdx.x == 0 && T.best) {
        T.best[0] = T.H ? bestH : ~0u;
        T.best[1] = bestC;
    }
    if (!T.mask) return;
    float F[9];
    for (int q = 0; q < 9; q++) F[q] = sFs[bm][q];
    for (uint32_t i = threadIdx.x; i < T.nHost; i += kThreads)
        T.mask[i] = (nm > 0 && i < n) ? (uint8_t)fgeom::inlier<kMode>(F, T.pts[i], t2) : 0;
}

// ---------------------------------------------------------------------------
// device sampler: hypothesis h of stream `seed` draws 8 distinct indices in
// [0, n) from a splitmix64 sequence seeded by (seed, h), rejecting repeats
__device__ inline uint64_t splitmix(uint64_t& s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr int kSampleMaxTasks = 32;
typedef ByValue<SampleTask, kSampleMaxTasks> SampleBatch;

__global__ void sample_kernel(uint32_t* idx, uint32_t H, uint32_t nHost, const uint32_t* nDev, uint64_t seed);

template <class Src>
__global__ void sample_batch_kernel(const Src batch)
{
    const SampleTask& T = batch[blockIdx.y];
    if (blockIdx.x * blockDim.x >= T.H) return;
    // same draw as sample_kernel, task taken from the batch
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= T.H) return;
    const uint32_t n = T.nDev ? min(*T.nDev, T.nHost) : T.nHost;
    const uint32_t k = T.k;
    uint32_t* out = T.idx + (size_t)h * k;
    if (n < k) {
        for (uint32_t i = 0; i < k; i++) out[i] = ~0u;
        return;
    }
    uint64_t s = T.seed ^ ((uint64_t)h * 0xD1B54A32D192ED03ull);
    uint32_t v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        v[i] = ~0u;
        if ((uint32_t)i >= k) continue;
        for (;;) {
            const uint32_t c = (uint32_t)(splitmix(s) % n);
            bool dup = false;
#pragma unroll
            for (int j = 0; j < i; j++) dup |= v[j] == c;
            if (!dup) {
                v[i] = c;
                break;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++)
        if ((uint32_t)i < k) out[i] = v[i];
}

// ---- weighted sampler for learned proposals (no reference counterpart; the
// reference draws uniformly, ransac.hpp:23-82). Weights become integers
// q_i = round(clamp(w_i, 0, 1) * 65536) with an exact u64 inclusive prefix, so
// the draw is reproducible bit for bit on the CPU. Hypothesis h uses the same
// splitmix64 stream as the uniform sampler: r = next % total selects the first
// i with cdf[i] > r (probability q_i / total); repeats are rejected (draws
// without replacement), at most kMaxWeightedDraws draws per hypothesis, after
// which it is marked ~0u. total == 0 falls back to the uniform draw.
constexpr uint32_t kMaxWeightedDraws = 512;

struct WSampleTask {
    uint32_t* idx;
    uint32_t H;
    uint32_t nHost;
    const uint32_t* nDev;
    const float* w;
    uint64_t* cdf;
    uint64_t seed;
};
struct WSampleBatch {
    WSampleTask t[kSampleMaxTasks];
};

__device__ inline uint32_t quant_weight(float w) { return (uint32_t)(fminf(fmaxf(w, 0.f), 1.f) * 65536.f + 0.5f); }

// one 1024-thread workgroup per task: cdf[i] = sum_{k <= i} q_k
__global__ __launch_bounds__(1024) void weight_scan_kernel(const WSampleBatch batch)
{
    const WSampleTask& T = batch.t[blockIdx.x];
    const uint32_t n = T.nDev ? min(*T.nDev, T.nHost) : T.nHost;
    __shared__ uint64_t sWave[16];
    __shared__ uint64_t sCarry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) sCarry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += 4096) {
        // each thread owns 4 consecutive entries
        const uint32_t i0 = base + 4 * tid;
        uint64_t q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) q[k] = i0 + k < n ? quant_weight(T.w[i0 + k]) : 0;
        q[1] += q[0];
        q[2] += q[1];
        q[3] += q[2];
        // inclusive scan of the thread totals across the wave
        uint64_t x = q[3];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) sWave[wave] = x;
        __syncthreads();
        uint64_t waveOff = 0;
        for (int w = 0; w < wave; w++) waveOff += sWave[w];
        const uint64_t off = sCarry + waveOff + x - q[3];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i0 + k < n) T.cdf[i0 + k] = off + q[k];
        __syncthreads();
        if (tid == 1023) sCarry = off + q[3];
        __syncthreads();
    }
}

__global__ void weighted_sample_kernel(const WSampleBatch batch)
{
    const WSampleTask& T = batch.t[blockIdx.y];
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= T.H) return;
    const uint32_t n = T.nDev ? min(*T.nDev, T.nHost) : T.nHost;
    uint32_t* out = T.idx + (size_t)h * 8;
    if (n < 8) {
        for (int i = 0; i < 8; i++) out[i] = ~0u;
        return;
    }
    const uint64_t total = T.cdf[n - 1];
    uint64_t s = T.seed ^ ((uint64_t)h * 0xD1B54A32D192ED03ull);
    uint32_t v[8];
    uint32_t draws = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        for (;;) {
            uint32_t c;
            if (total == 0) {
                c = (uint32_t)(splitmix(s) % n);
            } else {
                const uint64_t r = splitmix(s) % total;
                // first index with cdf > r
                uint32_t lo = 0, hi = n - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (T.cdf[mid] > r)
                        hi = mid;
                    else
                        lo = mid + 1;
                }
                c = lo;
            }
            bool dup = false;
#pragma unroll
            for (int j = 0; j < i; j++) dup |= v[j] == c;
            if (!dup) {
                v[i] = c;
                break;
            }
            if (++draws >= kMaxWeightedDraws) {
                for (int k = 0; k < 8; k++) out[k] = ~0u;
                return;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = v[i];
}

__global__ void sample_kernel(uint32_t* idx, uint32_t H, uint32_t nHost, const uint32_t* nDev, uint64_t seed)
{
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    const uint32_t n = nDev ? min(*nDev, nHost) : nHost;
    uint32_t* out = idx + (size_t)h * 8;
    if (n < 8) {
        for (int i = 0; i < 8; i++) out[i] = ~0u;
        return;
    }
    uint64_t s = seed ^ ((uint64_t)h * 0xD1B54A32D192ED03ull);
    uint32_t v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        for (;;) {
            const uint32_t c = (uint32_t)(splitmix(s) % n);
            bool dup = false;
#pragma unroll
            for (int j = 0; j < i; j++) dup |= v[j] == c;
            if (!dup) {
                v[i] = c;
                break;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = v[i];
}

int launch_sample_table(const SampleTask* tasks, uint32_t nTasks, uint32_t maxH, hipStream_t stream)
{
    if (nTasks == 0 || maxH == 0) return RSFM_OK;
    sample_batch_kernel<<<dim3(div_up(maxH, 256), nTasks), 256, 0, stream>>>(InTable<SampleTask>{tasks});
    return check_launch("sample table");
}

int launch_score_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxH, float threshold, int mode,
                       float minSepSq0, float minSepSq1, hipStream_t stream, bool withBest)
{
    RSFM_RETURN_IF(mode != 0 && mode != 1, RSFM_ERR_INVALID, "ransac: mode must be 0 or 1");
    if (nTasks == 0) return RSFM_OK;
    const float t2 = threshold * threshold;
    const InTable<DevTask> src{tasks};
    if (maxH > 0) {
        const dim3 grid(div_up(maxH, kScoreThreads), nTasks);
        if (mode == 0) score_kernel<0><<<grid, kScoreBlock, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
        else score_kernel<1><<<grid, kScoreBlock, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
        RSFM_CHECK(check_launch("ransac score table"));
    }
    if (!withBest) return RSFM_OK;
    if (mode == 0) best_kernel<0><<<nTasks, kBestThreads, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
    else best_kernel<1><<<nTasks, kBestThreads, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
    return check_launch("ransac best table");
}

int launch_score7_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxH, float threshold, int mode,
                        float minSepSq0, float minSepSq1, hipStream_t stream)
{
    RSFM_RETURN_IF(mode != 0 && mode != 1, RSFM_ERR_INVALID, "ransac7: mode must be 0 or 1");
    if (nTasks == 0) return RSFM_OK;
    const float t2 = threshold * threshold;
    const InTable<DevTask> src{tasks};
    if (maxH > 0) {
        const dim3 grid(div_up(maxH, kS7Samples), nTasks);
        if (mode == 0) score7_kernel<0><<<grid, kScoreBlock, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
        else score7_kernel<1><<<grid, kScoreBlock, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
        RSFM_CHECK(check_launch("ransac7 score table"));
    }
    if (mode == 0) best7_kernel<0><<<nTasks, kThreads, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
    else best7_kernel<1><<<nTasks, kThreads, 0, stream>>>(src, t2, minSepSq0, minSepSq1);
    return check_launch("ransac7 best table");
}

}  // namespace ransac
}  // namespace rsfm_amd

using namespace rsfm_amd;

namespace rsfm_amd {
namespace ransac {
// score + best over batches of at most kMaxTasks tasks; kind = model kind
// (0 / 1 fundamental matrix, 2 homography)
static int run_hypotheses(const rsfm_ransac_task* tasks, size_t nbTasks, float threshold, int kind, float minSepSq0,
                          float minSepSq1, hipStream_t stream, const char* name)
{
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, std::string(name) + ": tasks is null");
    RSFM_RETURN_IF(!(threshold > 0.f), RSFM_ERR_INVALID, std::string(name) + ": threshold must be > 0");
    const float t2 = threshold * threshold;
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        Batch batch{};
        uint32_t maxH = 0;
        for (size_t i = 0; i < n; i++) {
            const rsfm_ransac_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.nbHyp > 0 && (t.sampleIdx == nullptr || t.counts == nullptr), RSFM_ERR_INVALID,
                           std::string(name) + ": sampleIdx/counts must be set when nbHyp > 0");
            RSFM_RETURN_IF(t.nbPts > 0 && t.pts == nullptr, RSFM_ERR_INVALID, std::string(name) + ": pts is null");
            RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16, RSFM_ERR_INVALID,
                           std::string(name) + ": pts must be 16-byte aligned");
            batch.t[i] = DevTask{reinterpret_cast<const float4*>(t.pts), t.nbPtsDev, t.nbPts, t.nbHyp, t.sampleIdx,
                                 t.counts, t.models, t.best, t.inlierMask};
            maxH = std::max(maxH, t.nbHyp);
        }
        if (maxH > 0) {
            const dim3 grid(div_up(maxH, kScoreThreads), (uint32_t)n);
            if (kind == 0) score_kernel<0><<<grid, kScoreBlock, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            else if (kind == 1) score_kernel<1><<<grid, kScoreBlock, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            else score_kernel<2><<<grid, kScoreBlock, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            RSFM_CHECK(check_launch(name));
        }
        bool anyBest = false;
        for (size_t i = 0; i < n; i++) anyBest |= batch.t[i].best != nullptr || batch.t[i].mask != nullptr;
        if (anyBest) {
            if (kind == 0) best_kernel<0><<<(uint32_t)n, kBestThreads, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            else if (kind == 1) best_kernel<1><<<(uint32_t)n, kBestThreads, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            else best_kernel<2><<<(uint32_t)n, kBestThreads, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            RSFM_CHECK(check_launch(name));
        }
    }
    return RSFM_OK;
}
}  // namespace ransac
}  // namespace rsfm_amd

extern "C" int rsfm_ransac_f8(const rsfm_ransac_task* tasks, size_t nbTasks, float threshold, int mode,
                              float minSepSq0, float minSepSq1, hipStream_t stream)
{
    RSFM_RETURN_IF(mode != 0 && mode != 1, RSFM_ERR_INVALID, "ransac: mode must be 0 (two-sided line) or 1 (Sampson)");
    return ransac::run_hypotheses(tasks, nbTasks, threshold, mode, minSepSq0, minSepSq1, stream, "ransac");
}

extern "C" int rsfm_ransac_h4(const rsfm_ransac_task* tasks, size_t nbTasks, float threshold, float minSepSq0,
                              float minSepSq1, hipStream_t stream)
{
    return ransac::run_hypotheses(tasks, nbTasks, threshold, 2, minSepSq0, minSepSq1, stream, "ransac_h4");
}

extern "C" int rsfm_ransac_sample_indices(uint32_t* sampleIdx, uint32_t nbHyp, uint32_t nbPts,
                                          const uint32_t* nbPtsDev, uint64_t seed, hipStream_t stream)
{
    if (nbHyp == 0) return RSFM_OK;
    RSFM_RETURN_IF(sampleIdx == nullptr, RSFM_ERR_INVALID, "sample_indices: sampleIdx is null");
    ransac::sample_kernel<<<div_up(nbHyp, 256), 256, 0, stream>>>(sampleIdx, nbHyp, nbPts, nbPtsDev, seed);
    return check_launch("ransac sample_indices");
}

namespace {
int sample_batch(const rsfm_sample_task* tasks, size_t nbTasks, uint32_t k, hipStream_t stream)
{
    using namespace ransac;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "sample batch: tasks is null");
    RSFM_RETURN_IF(k < 4 || k > 8, RSFM_ERR_INVALID, "sample batch: sampleSize must be in [4, 8]");
    for (size_t b = 0; b < nbTasks; b += kSampleMaxTasks) {
        const size_t n = std::min<size_t>(kSampleMaxTasks, nbTasks - b);
        SampleBatch batch{};
        uint32_t maxH = 0;
        for (size_t i = 0; i < n; i++) {
            const rsfm_sample_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.nbHyp > 0 && t.sampleIdx == nullptr, RSFM_ERR_INVALID, "sample batch: sampleIdx is null");
            batch.t[i] = SampleTask{t.sampleIdx, t.nbHyp, t.nbPts, t.nbPtsDev, t.seed, k};
            maxH = std::max(maxH, t.nbHyp);
        }
        if (maxH == 0) continue;
        sample_batch_kernel<<<dim3(div_up(maxH, 256), (uint32_t)n), 256, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("ransac sample_indices_batch"));
    }
    return RSFM_OK;
}
}  // namespace

extern "C" int rsfm_ransac_sample_indices_batch(const rsfm_sample_task* tasks, size_t nbTasks, hipStream_t stream)
{
    return sample_batch(tasks, nbTasks, 8, stream);
}

extern "C" int rsfm_ransac_sample_indices_batch_k(const rsfm_sample_task* tasks, size_t nbTasks, uint32_t sampleSize,
                                                  hipStream_t stream)
{
    return sample_batch(tasks, nbTasks, sampleSize, stream);
}

extern "C" int rsfm_ransac_sample_weighted_batch(const rsfm_weighted_sample_task* tasks, size_t nbTasks,
                                                 hipStream_t stream)
{
    using namespace ransac;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "weighted sample: tasks is null");
    for (size_t b = 0; b < nbTasks; b += kSampleMaxTasks) {
        const size_t n = std::min<size_t>(kSampleMaxTasks, nbTasks - b);
        WSampleBatch batch{};
        uint32_t maxH = 0;
        for (size_t i = 0; i < n; i++) {
            const rsfm_weighted_sample_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.nbHyp > 0 && t.sampleIdx == nullptr, RSFM_ERR_INVALID,
                           "weighted sample: sampleIdx is null");
            RSFM_RETURN_IF(t.nbPts > 0 && (t.weights == nullptr || t.cdf == nullptr), RSFM_ERR_INVALID,
                           "weighted sample: weights/cdf null");
            RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.cdf) % 8 != 0, RSFM_ERR_INVALID,
                           "weighted sample: cdf must be 8-byte aligned");
            batch.t[i] = WSampleTask{t.sampleIdx, t.nbHyp, t.nbPts, t.nbPtsDev, t.weights, t.cdf, t.seed};
            maxH = std::max(maxH, t.nbHyp);
        }
        weight_scan_kernel<<<(uint32_t)n, 1024, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("weighted sample scan"));
        if (maxH == 0) continue;
        weighted_sample_kernel<<<dim3(div_up(maxH, 256), (uint32_t)n), 256, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("weighted sample"));
    }
    return RSFM_OK;
}

extern "C" int rsfm_ransac_f7(const rsfm_ransac_task* tasks, size_t nbTasks, float threshold, int mode, float minSepSq0,
                              float minSepSq1, hipStream_t stream)
{
    using namespace ransac;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "ransac7: tasks is null");
    RSFM_RETURN_IF(mode != 0 && mode != 1, RSFM_ERR_INVALID, "ransac7: mode must be 0 (two-sided line) or 1 (Sampson)");
    RSFM_RETURN_IF(!(threshold > 0.f), RSFM_ERR_INVALID, "ransac7: threshold must be > 0");
    const float t2 = threshold * threshold;
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        Batch batch{};
        uint32_t maxH = 0;
        for (size_t i = 0; i < n; i++) {
            const rsfm_ransac_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.nbHyp > 0 && (t.sampleIdx == nullptr || t.counts == nullptr), RSFM_ERR_INVALID,
                           "ransac7: sampleIdx/counts must be set when nbHyp > 0");
            RSFM_RETURN_IF(t.nbPts > 0 && t.pts == nullptr, RSFM_ERR_INVALID, "ransac7: pts is null");
            RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16, RSFM_ERR_INVALID, "ransac7: pts must be 16-byte aligned");
            batch.t[i] = DevTask{reinterpret_cast<const float4*>(t.pts), t.nbPtsDev, t.nbPts, t.nbHyp, t.sampleIdx,
                                 t.counts, t.models, t.best, t.inlierMask};
            maxH = std::max(maxH, t.nbHyp);
        }
        if (maxH > 0) {
            const dim3 grid(div_up(maxH, kS7Samples), (uint32_t)n);
            if (mode == 0) score7_kernel<0><<<grid, kScoreBlock, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            else score7_kernel<1><<<grid, kScoreBlock, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            RSFM_CHECK(check_launch("ransac7 score"));
        }
        bool anyBest = false;
        for (size_t i = 0; i < n; i++) anyBest |= batch.t[i].best != nullptr || batch.t[i].mask != nullptr;
        if (anyBest) {
            if (mode == 0) best7_kernel<0><<<(uint32_t)n, kThreads, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            else best7_kernel<1><<<(uint32_t)n, kThreads, 0, stream>>>(batch, t2, minSepSq0, minSepSq1);
            RSFM_CHECK(check_launch("ransac7 best"));
        }
    }
    return RSFM_OK;
}
