// ransac_finish.hip - the end of the 8-point F-RANSAC driver: adaptive cut of
// a scored hypothesis table and the least-squares refit on the inliers.
//
// Reference path: ransac<F,8> (ransac.hpp:115-163) stops after
// min(maxIter, ceil(log2(1 - confidence) / log2(1 - (best / n)^8))) sane draws
// (recomputed at every strictly better draw) or after more than
// min(128, maxIter) insane draws in a row; then
// findEpipolarityRansacWithInsufficientTiePts<0> (ImagePair.cpp:185-203)
// refits F on the best sample, counts its inliers and, with at least 8 of
// them, refits F on all inliers and keeps that when it has more inliers.
//
// GPU layout: one 256-thread workgroup per pair.
//  * Cut: the table is walked in chunks of 1024 hypotheses (4 per thread);
//    per chunk one block scan gives every hypothesis its sane-draw count, the
//    best key so far (count << 32 | ~h) and the last sane index before it, so
//    each hypothesis tests the reference's two stop conditions on its own;
//    the first one that stops ends the walk.
//  * Refit: three passes over the points, each thread taking points t,
//    t + 256, ... in order (inlier test of the best F, then FP64 sums):
//    centroids, mean distances to them (Hartley scales), the 45 entries of
//    the normal matrix of the normalised rows. An entry's 256 thread
//    partials are summed in a fixed two-level order (16 groups of the
//    threads congruent mod 16, each in thread order, then the group sums in
//    group order), so the CPU restatement repeats every rounding with a
//    dependency chain of 32 additions instead of 256. The null vector,
//    rank-2 projection and denormalisation are the FP64 single-lane
//    routines of hgeom.hpp.
#include "internal.hpp"
#include "fgeom.hpp"
#include "hgeom.hpp"
#include <algorithm>

namespace rsfm_amd {
namespace ransac {

constexpr int kFinThreads = 256;
constexpr int kFinWaves = kFinThreads / 64;
constexpr int kFinPer = 4;
constexpr uint32_t kFinChunk = kFinThreads * kFinPer;
// pairs per launch of the C-ABI entry point: one workgroup each, so a launch
// lasts about as long as its slowest pair's refit (a chain of FP64 steps);
// 32 tasks of 88 bytes are 2.8 KiB of kernel arguments
constexpr int kFinMaxTasks = 32;
constexpr int kNormEntries = 45;  // upper triangle of the 9x9 normal matrix

// log2 of x in (0, 1] (0 -> -inf, negative -> NaN): x = m 2^e with m in
// [sqrt(1/2), sqrt(2)], ln m = 2 atanh(s) with s = (m - 1) / (m + 1) summed as
// s (2 + z P(z)), z = s^2 (|s| <= 0.1716, terms to z^11). Round-to-nearest
// FP64 operations only, in a fixed order.
__host__ __device__ inline double det_log2(double x)
{
    if (x == 1.0) return 0.0;
    if (x == 0.0) return -__builtin_inf();
    if (!(x > 0.0)) return __builtin_nan("");  // negative or NaN, as libm
    uint64_t bits;
    __builtin_memcpy(&bits, &x, 8);
    int e = (int)((bits >> 52) & 0x7FF) - 1023;
    bits = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    __builtin_memcpy(&m, &bits, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 2.0 / 23.0;
    p = p * z + 2.0 / 21.0;
    p = p * z + 2.0 / 19.0;
    p = p * z + 2.0 / 17.0;
    p = p * z + 2.0 / 15.0;
    p = p * z + 2.0 / 13.0;
    p = p * z + 2.0 / 11.0;
    p = p * z + 2.0 / 9.0;
    p = p * z + 2.0 / 7.0;
    p = p * z + 2.0 / 5.0;
    p = p * z + 2.0 / 3.0;
    const double lnm = s * (2.0 + z * p);
    return (double)e + lnm * 1.4426950408889634;
}

// ransac.hpp:155-157 with det_log2 and (best / n)^8 by squaring; num =
// det_log2(1 - confidence). The conversions of a non-finite quotient follow
// the reference built for x86-64: -inf and NaN give maxIter, +inf gives 0.
__host__ __device__ inline uint32_t required_iters(uint32_t best, uint32_t n, double num, uint32_t maxIter)
{
    const double p = (double)best / (double)n;
    const double p2 = p * p;
    const double p4 = p2 * p2;
    const double p8 = p4 * p4;
    const double q = num / det_log2(1.0 - p8);
    if (q != q || q == -__builtin_inf()) return maxIter;
    if (q == __builtin_inf()) return 0u;
    const double c = __builtin_ceil(q);
    if (!(c > 0.0)) return 0u;
    if (c >= (double)maxIter) return maxIter;
    return (uint32_t)c;
}

typedef ByValue<FinishTask, kFinMaxTasks> FinishBatch;
static_assert(sizeof(FinishBatch) + sizeof(FinishParams) <= 4096, "the batch is kept within 4 KiB of kernel arguments");

constexpr int kSumGroups = 16;  // kFinThreads = kSumGroups * kSumGroups
static_assert(kSumGroups * kSumGroups == kFinThreads, "two-level sums of the thread partials");

struct FinishSmem {
    double part[kNormEntries][kFinThreads];
    double grp[kNormEntries][kSumGroups];
    double tot[kNormEntries];
    uint64_t waveKey[kFinWaves];
    uint32_t waveCnt[kFinWaves];
    int32_t waveLast[kFinWaves];
    uint32_t firstStop;
    uint32_t stop;
    uint64_t stopKey;
    float F[9];
    float G[9];
    int ok;
    hgeom::Norm nm;
};

template <class T, class Op>
__device__ inline T wave_scan_incl(T x, Op op)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, o);
        if (lane >= o) x = op(x, y);
    }
    return x;
}

// sum of the workgroup's per-thread counts (integers: any order is exact)
__device__ inline uint32_t block_sum_u32(FinishSmem& s, uint32_t c)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o);
    if ((threadIdx.x & 63) == 0) s.waveCnt[threadIdx.x >> 6] = c;
    __syncthreads();
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kFinWaves; w++) total += s.waveCnt[w];
    __syncthreads();
    return total;
}

// fn(p) for each inlier p of model F among this thread's points t, t + 256,
// ... in index order; the loads of four points are issued before any test
constexpr int kFinUnroll = 4;
template <int kMode, class Fn>
__device__ inline void for_inliers(const float F[9], const float4* pts, uint32_t n, float t2, Fn fn)
{
    for (uint32_t i0 = threadIdx.x; i0 < n; i0 += kFinUnroll * kFinThreads) {
        float4 p[kFinUnroll];
#pragma unroll
        for (int u = 0; u < kFinUnroll; u++) {
            const uint32_t i = i0 + u * kFinThreads;
            p[u] = i < n ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < kFinUnroll; u++)
            if (i0 + u * kFinThreads < n && fgeom::inlier<kMode>(F, p[u], t2)) fn(p[u]);
    }
}

// inliers of model F (ok) among the n points
template <int kMode>
__device__ uint32_t count_model(FinishSmem& s, const float F[9], bool ok, const float4* pts, uint32_t n, float t2)
{
    uint32_t c = 0;
    if (ok) for_inliers<kMode>(F, pts, n, t2, [&](const float4&) { c++; });
    return block_sum_u32(s, c);
}

// the first `entries` rows of s.part summed into s.tot: group g adds the
// partials of threads g, g + 16, ..., g + 240 in that order, then the 16
// group sums are added in group order
__device__ inline void sum_parts(FinishSmem& s, int entries)
{
    __syncthreads();
    for (int task = threadIdx.x; task < entries * kSumGroups; task += kFinThreads) {
        const int e = task / kSumGroups, g = task % kSumGroups;
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < kSumGroups; k++) a = a + s.part[e][k * kSumGroups + g];
        s.grp[e][g] = a;
    }
    __syncthreads();
    if ((int)threadIdx.x < entries) {
        double a = 0.0;
#pragma unroll
        for (int g = 0; g < kSumGroups; g++) a = a + s.grp[threadIdx.x][g];
        s.tot[threadIdx.x] = a;
    }
    __syncthreads();
}

// least-squares F on the m inliers of s.F; result in s.G, s.ok (of the refit)
template <int kMode>
__device__ bool refit_f(FinishSmem& s, const float4* pts, uint32_t n, float t2, uint32_t m)
{
    float F[9];
#pragma unroll
    for (int q = 0; q < 9; q++) F[q] = s.F[q];
    const uint32_t t = threadIdx.x;
    // centroids
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for_inliers<kMode>(F, pts, n, t2, [&](const float4& p) {
        a[0] = a[0] + (double)p.x;
        a[1] = a[1] + (double)p.y;
        a[2] = a[2] + (double)p.z;
        a[3] = a[3] + (double)p.w;
    });
#pragma unroll
    for (int k = 0; k < 4; k++) s.part[k][t] = a[k];
    sum_parts(s, 4);
    const double dm = (double)m;
    double c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = s.tot[k] / dm;
    // mean distances to the centroids
    double d0 = 0.0, d1 = 0.0;
    for_inliers<kMode>(F, pts, n, t2, [&](const float4& p) {
        const double dx0 = (double)p.x - c[0], dy0 = (double)p.y - c[1];
        const double dx1 = (double)p.z - c[2], dy1 = (double)p.w - c[3];
        d0 = d0 + sqrt(dx0 * dx0 + dy0 * dy0);
        d1 = d1 + sqrt(dx1 * dx1 + dy1 * dy1);
    });
    s.part[0][t] = d0;
    s.part[1][t] = d1;
    sum_parts(s, 2);
    const double n0 = s.tot[0] / dm, n1 = s.tot[1] / dm;
    if (!(n0 > 0.0) || !(n1 > 0.0)) return false;
    const double s0 = hgeom::kSqrt2 / n0, s1 = hgeom::kSqrt2 / n1;
    // normal matrix of the rows [x1 x0, y1 x0, x0, x1 y0, y1 y0, y0, x1, y1, 1]
    double acc[kNormEntries];
#pragma unroll
    for (int e = 0; e < kNormEntries; e++) acc[e] = 0.0;
    for_inliers<kMode>(F, pts, n, t2, [&](const float4& p) {
        const double x0 = ((double)p.x - c[0]) * s0, y0 = ((double)p.y - c[1]) * s0;
        const double x1 = ((double)p.z - c[2]) * s1, y1 = ((double)p.w - c[3]) * s1;
        const double r[9] = {x1 * x0, y1 * x0, x0, x1 * y0, y1 * y0, y0, x1, y1, 1.0};
        int e = 0;
#pragma unroll
        for (int u = 0; u < 9; u++)
#pragma unroll
            for (int v = u; v < 9; v++) {
                acc[e] = acc[e] + r[u] * r[v];
                e++;
            }
    });
#pragma unroll
    for (int e = 0; e < kNormEntries; e++) s.part[e][t] = acc[e];
    sum_parts(s, kNormEntries);
    if (t == 0) {
        double N[9][9];
        int e = 0;
        for (int u = 0; u < 9; u++)
            for (int v = u; v < 9; v++) {
                N[u][v] = s.tot[e];
                N[v][u] = s.tot[e];
                e++;
            }
        hgeom::Norm nm;
#pragma unroll
        for (int k = 0; k < 4; k++) nm.c[k] = c[k];
        nm.s0 = s0;
        nm.s1 = s1;
        double v[9];
        bool ok = hgeom::null_vector9(N, v);
        if (ok) ok = hgeom::finish_f(v, nm, s.G);
        s.ok = ok;
    }
    __syncthreads();
    const bool ok = s.ok != 0;
    __syncthreads();
    return ok;
}

// the walk of ransac.hpp:139-160 over counts[0, H): s.stop = hypotheses
// walked, s.stopKey = best key among them (0 when none was sane)
__device__ void adaptive_cut(FinishSmem& s, const uint32_t* counts, uint32_t H, uint32_t n, float conf,
                             uint32_t maxIter)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double num = det_log2(1.0 - (double)conf);
    const uint32_t cap = min(maxIter, 128u);
    uint32_t itersBefore = 0;
    uint64_t keyBefore = 0;
    int32_t lastBefore = -1;
    if (threadIdx.x == 0) {
        s.stop = H;
        s.stopKey = 0;
    }
    for (uint32_t base = 0; base < H; base += kFinChunk) {
        const uint32_t h0 = base + threadIdx.x * kFinPer;
        uint32_t c[kFinPer];
        uint32_t lc = 0;
        uint64_t lk = 0;
        int32_t ll = -1;
#pragma unroll
        for (int j = 0; j < kFinPer; j++) {
            const uint32_t h = h0 + j;
            c[j] = h < H ? counts[h] : 0u;
            if (c[j] > 0) {
                lc++;
                lk = max(lk, ((uint64_t)c[j] << 32) | (uint32_t)~h);
                ll = (int32_t)h;
            }
        }
        // inclusive wave scans, then the wave totals through LDS
        const uint32_t ic = wave_scan_incl(lc, [](uint32_t x, uint32_t y) { return x + y; });
        const uint64_t ik = wave_scan_incl(lk, [](uint64_t x, uint64_t y) { return x > y ? x : y; });
        const int32_t il = wave_scan_incl(ll, [](int32_t x, int32_t y) { return x > y ? x : y; });
        if (lane == 63) {
            s.waveCnt[wave] = ic;
            s.waveKey[wave] = ik;
            s.waveLast[wave] = il;
        }
        __syncthreads();
        uint32_t it = itersBefore;
        uint64_t key = keyBefore;
        int32_t last = lastBefore;
        for (int w = 0; w < wave; w++) {
            it += s.waveCnt[w];
            key = max(key, s.waveKey[w]);
            last = max(last, s.waveLast[w]);
        }
        uint32_t totCnt = 0;
        uint64_t totKey = keyBefore;
        int32_t totLast = lastBefore;
        for (int w = 0; w < kFinWaves; w++) {
            totCnt += s.waveCnt[w];
            totKey = max(totKey, s.waveKey[w]);
            totLast = max(totLast, s.waveLast[w]);
        }
        // exclusive within the wave
        it += ic - lc;
        const uint64_t ek = __shfl_up(ik, 1);
        const int32_t el = __shfl_up(il, 1);
        if (lane > 0) {
            key = max(key, ek);
            last = max(last, el);
        }
        uint32_t myStop = ~0u;
        uint64_t myKey = 0;
#pragma unroll
        for (int j = 0; j < kFinPer; j++) {
            const uint32_t h = h0 + j;
            if (h >= H || myStop != ~0u) continue;
            if (c[j] > 0) {
                it++;
                key = max(key, ((uint64_t)c[j] << 32) | (uint32_t)~h);
                last = (int32_t)h;
                if (it >= required_iters((uint32_t)(key >> 32), n, num, maxIter)) {
                    myStop = h;
                    myKey = key;
                }
            } else {
                const uint32_t run = h - (uint32_t)(last + 1) + 1u;  // insane draws in a row, this one included
                if (run > cap) {
                    myStop = h;
                    myKey = key;
                }
            }
        }
        __syncthreads();  // wave totals read by everyone
        if (threadIdx.x == 0) s.firstStop = ~0u;
        __syncthreads();
        if (myStop != ~0u) atomicMin(&s.firstStop, myStop);
        __syncthreads();
        const uint32_t first = s.firstStop;
        if (first != ~0u) {
            if (myStop == first) {
                s.stop = first + 1;
                s.stopKey = myKey;
            }
            __syncthreads();
            return;
        }
        itersBefore += totCnt;
        keyBefore = totKey;
        lastBefore = totLast;
        __syncthreads();
    }
    if (threadIdx.x == 0) s.stopKey = keyBefore;
    __syncthreads();
}

template <int kMode, class Src>
__global__ __launch_bounds__(kFinThreads) void finish_kernel(const Src batch, const FinishParams prm)
{
    __shared__ FinishSmem s;
    const FinishTask& T = batch[blockIdx.x];
    const uint32_t n = T.nDev ? min(*T.nDev, T.nHost) : T.nHost;
    const uint32_t t = threadIdx.x;
    const float t2 = prm.t2;
    if (n < 8) {
        for (uint32_t i = t; i < T.nHost; i += kFinThreads)
            if (T.mask) T.mask[i] = 0;
        if (t < 9) T.F[t] = 0.f;
        if (t == 0) {
            T.result[0] = ~0u;
            T.result[1] = 0u;
            T.result[2] = 0u;
            T.result[3] = 0u;
            if (T.best) {
                T.best[0] = ~0u;
                T.best[1] = 0u;
            }
        }
        return;
    }
    // ---- selection
    if (prm.conf > 0.f) {
        adaptive_cut(s, T.counts, T.H, n, prm.conf, prm.maxIter);
    } else {
        // keys are distinct, so their maximum does not depend on the order;
        // four counts are loaded per trip before they are compared
        uint64_t key = 0;
        for (uint32_t h0 = t; h0 < T.H; h0 += kFinUnroll * kFinThreads) {
            uint32_t cnt[kFinUnroll];
#pragma unroll
            for (int u = 0; u < kFinUnroll; u++) {
                const uint32_t h = h0 + u * kFinThreads;
                cnt[u] = h < T.H ? T.counts[h] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kFinUnroll; u++)
                if (cnt[u] > 0) key = max(key, ((uint64_t)cnt[u] << 32) | (uint32_t)~(h0 + u * kFinThreads));
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint64_t k = __shfl_xor(key, o);
            key = k > key ? k : key;
        }
        if ((t & 63) == 0) s.waveKey[t >> 6] = key;
        __syncthreads();
        if (t == 0) {
            uint64_t k = 0;
            for (int w = 0; w < kFinWaves; w++) k = max(k, s.waveKey[w]);
            s.stopKey = k;
            s.stop = T.H;
        }
        __syncthreads();
    }
    const uint64_t bk = s.stopKey;
    const uint32_t walked = s.stop;
    const uint32_t bestH = bk ? ~(uint32_t)bk : ~0u;
    // ---- the best sample's F (the reference's fallback sample 0..7 when none was sane)
    if (t == 0) {
        float F[9];
        bool ok;
        if (bestH != ~0u && T.models != nullptr) {
            for (int q = 0; q < 9; q++) F[q] = T.models[(size_t)bestH * 9 + q];
            ok = true;
        } else {
            float4 p[8];
            for (int i = 0; i < 8; i++) {
                const uint32_t k = bestH != ~0u ? T.idx[(size_t)bestH * 8 + i] : (uint32_t)i;
                p[i] = T.pts[k < n ? k : 0];
            }
            ok = fgeom::fit8(p, F);
        }
        for (int q = 0; q < 9; q++) s.F[q] = ok ? F[q] : 0.f;
        s.ok = ok;
    }
    __syncthreads();
    uint32_t c0;
    {
        float F[9];
#pragma unroll
        for (int q = 0; q < 9; q++) F[q] = s.F[q];
        c0 = count_model<kMode>(s, F, s.ok != 0, T.pts, n, t2);
    }
    uint32_t cFinal = c0;
    // ---- refit on the inliers
    if (prm.refit && c0 >= 8 && s.ok) {
        if (refit_f<kMode>(s, T.pts, n, t2, c0)) {
            float G[9];
#pragma unroll
            for (int q = 0; q < 9; q++) G[q] = s.G[q];
            const uint32_t c1 = count_model<kMode>(s, G, true, T.pts, n, t2);
            if (c1 > c0) {
                cFinal = c1;
                if (t < 9) s.F[t] = G[t];
            }
            __syncthreads();
        }
    }
    // ---- outputs
    float F[9];
#pragma unroll
    for (int q = 0; q < 9; q++) F[q] = s.F[q];
    const bool ok = s.ok != 0;
    if (T.mask)
        for (uint32_t i0 = t; i0 < T.nHost; i0 += kFinUnroll * kFinThreads) {
            float4 p[kFinUnroll];
#pragma unroll
            for (int u = 0; u < kFinUnroll; u++) {
                const uint32_t i = i0 + u * kFinThreads;
                p[u] = (ok && i < n) ? T.pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < kFinUnroll; u++) {
                const uint32_t i = i0 + u * kFinThreads;
                if (i < T.nHost) T.mask[i] = (ok && i < n) ? (uint8_t)fgeom::inlier<kMode>(F, p[u], t2) : (uint8_t)0;
            }
        }
    if (t < 9) T.F[t] = F[t];
    if (t == 0) {
        T.result[0] = bestH;
        T.result[1] = c0;
        T.result[2] = walked;
        T.result[3] = cFinal;
        if (T.best) {
            T.best[0] = bestH;
            T.best[1] = bestH != ~0u ? (uint32_t)(bk >> 32) : 0u;
        }
    }
}

int make_finish_params(float threshold, int mode, float confidence, uint32_t maxIter, int refit, FinishParams& p)
{
    RSFM_RETURN_IF(mode != 0 && mode != 1, RSFM_ERR_INVALID, "ransac finish: mode must be 0 or 1");
    RSFM_RETURN_IF(!(threshold > 0.f), RSFM_ERR_INVALID, "ransac finish: threshold must be > 0");
    RSFM_RETURN_IF(!(confidence <= 1.f), RSFM_ERR_INVALID, "ransac finish: confidence must be <= 1");
    p = FinishParams{threshold * threshold, confidence, maxIter, refit};
    return RSFM_OK;
}

int launch_finish_table(const FinishTask* tasks, uint32_t nTasks, const FinishParams& p, int mode, hipStream_t stream)
{
    if (nTasks == 0) return RSFM_OK;
    const InTable<FinishTask> src{tasks};
    if (mode == 0) finish_kernel<0><<<nTasks, kFinThreads, 0, stream>>>(src, p);
    else finish_kernel<1><<<nTasks, kFinThreads, 0, stream>>>(src, p);
    return check_launch("ransac finish table");
}

}  // namespace ransac
}  // namespace rsfm_amd

using namespace rsfm_amd;

extern "C" int rsfm_ransac_f8_finish(const rsfm_ransac_finish_task* tasks, size_t nbTasks, float threshold, int mode,
                                     float confidence, uint32_t maxIter, int refit, hipStream_t stream)
{
    using namespace ransac;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "ransac finish: tasks is null");
    FinishParams prm;
    RSFM_CHECK(make_finish_params(threshold, mode, confidence, maxIter, refit, prm));
    for (size_t b = 0; b < nbTasks; b += kFinMaxTasks) {
        const size_t n = std::min<size_t>(kFinMaxTasks, nbTasks - b);
        FinishBatch batch{};
        for (size_t i = 0; i < n; i++) {
            const rsfm_ransac_finish_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.F == nullptr || t.result == nullptr, RSFM_ERR_INVALID, "ransac finish: F/result is null");
            RSFM_RETURN_IF(t.nbPts > 0 && t.pts == nullptr, RSFM_ERR_INVALID, "ransac finish: pts is null");
            RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16, RSFM_ERR_INVALID,
                           "ransac finish: pts must be 16-byte aligned");
            RSFM_RETURN_IF(t.nbHyp > 0 && (t.counts == nullptr || (t.models == nullptr && t.sampleIdx == nullptr)),
                           RSFM_ERR_INVALID, "ransac finish: counts and models or sampleIdx must be set");
            batch.t[i] = FinishTask{reinterpret_cast<const float4*>(t.pts), t.nbPtsDev, t.nbPts, t.sampleIdx, t.counts,
                                    t.models, t.nbHyp, t.F, t.result, t.inlierMask, nullptr};
        }
        if (mode == 0) finish_kernel<0><<<(uint32_t)n, kFinThreads, 0, stream>>>(batch, prm);
        else finish_kernel<1><<<(uint32_t)n, kFinThreads, 0, stream>>>(batch, prm);
        RSFM_CHECK(check_launch("ransac finish"));
    }
    return RSFM_OK;
}
