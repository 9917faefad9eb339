// degensac.hip - DEGENSAC finisher: local optimisation, the dominant-plane
// degeneracy test and the plane-and-parallax F search, one workgroup per pair.
//
// Reference path: findEpipolarityDegensac (degensac.cpp:102-240). The
// reference draws 7-point samples one at a time and, each time a sample's
// best F (after checkEpipolarityAndOptimise, degensac.cpp:65-78) beats the
// best so far, runs the degeneracy test on it (degensac.cpp:151-187), and for
// a sample with >= 5 points on one homography refines that H
// (checkHomographyAndOptimise, degensac.cpp:87-100) and searches F among the
// off-plane points by 2-point plane-and-parallax RANSAC
// (degensac.cpp:195-228).
//
// Here the 7-point samples are scored in bulk first (rsfm_ransac_f7), and this
// kernel replays the reference's per-new-best logic over the nbCandidates
// highest-scoring samples in descending order (count, then lowest index):
//  - each of a sample's (up to three) F is locally optimised: inlier count
//    n0; with n0 >= 8, a least-squares refit on at most 128 inliers (evenly
//    spaced by rank among the inliers, where the reference takes a seeded
//    Bernoulli subsample, degensac.cpp:33-55) and its count n1; the refit
//    replaces F when n1 > n0;
//  - a sample whose best optimised count beats the best so far becomes the
//    best and gets the degeneracy test: H from F and each of the five sample
//    triplets {0,1,2} {3,4,5} {0,1,6} {3,4,6} {2,5,6} (the table of the
//    DEGENSAC paper, also degensac.cpp:152-165), scored on the other four
//    sample points; 3 + most agreeing >= 5 marks the sample planar;
//  - a planar sample's H is optimised the same way (>= 4 inliers, refit on at
//    most 128) and kept when it has the most H inliers so far; the points off
//    that H feed nbPlaneSamples 2-point samples (splitmix64 stream, ranks into
//    the off-plane set), each giving F = [e']x H; the one with most inliers
//    among the off-plane points (lowest sample on ties) is optimised on all
//    points and replaces the best F when it has more inliers.
// The reference's adaptive iteration bound (degensac.cpp:230) is replaced by
// the caller's fixed hypothesis count.
//
// Everything that body computes for a candidate depends on that candidate
// alone; only whether it runs (the candidate beats the best so far) and what
// it replaces depend on the earlier ones. So three launches: the nbCandidates
// best samples of each pair (select_kernel), one workgroup per (pair,
// candidate) that runs the whole body for its candidate and stores a record
// (candidate_kernel), and one workgroup per pair that replays the records in
// order with the reference's acceptance rules and writes the results and
// masks (replay_kernel): the same outputs as one workgroup walking the
// candidates in turn, with nbCandidates times the parallelism. The
// plane-and-parallax search of a planar candidate that can still lead is
// itself split over workgroups of 16 samples (plane_kernel) and its winner
// optimised after (plane_finish_kernel).
//
// GPU layout: every per-point pass is spread over the 256 lanes (ballot
// inlier bits into LDS bitmaps, integer counts), the plane search keeps its
// samples' F in LDS and every wavefront scores all of them over a quarter of
// each LDS point tile, and the small dense solves run on one lane in FP64
// (hgeom.hpp).
#include "internal.hpp"
#include "fgeom.hpp"
#include "hgeom.hpp"
#include <algorithm>

namespace rsfm_amd {
namespace degensac {

constexpr int kThreads = 256;
constexpr uint32_t kMaxPts = 65536;
constexpr uint32_t kMaxWords = kMaxPts / 32;
constexpr int kMaxSel = 128;  // maxNbSamplesForFHOpt, degensac.cpp:31
constexpr uint32_t kMaxCand = 64;
constexpr uint32_t kMaxPlaneSamples = 4096;
constexpr int kTile = 1024;
constexpr int kMaxTasks = 16;

typedef ByValue<DevTask, kMaxTasks> Batch;

__device__ __constant__ int kTri[5][3] = {{0, 1, 2}, {3, 4, 5}, {0, 1, 6}, {3, 4, 6}, {2, 5, 6}};
__device__ __constant__ int kRem[5][4] = {{3, 4, 5, 6}, {0, 1, 2, 6}, {2, 3, 4, 5}, {0, 1, 2, 5}, {0, 1, 3, 4}};

struct FitScratch {
    double pt[kMaxSel][4];
    double d0[kMaxSel], d1[kMaxSel];
    double row[2 * kMaxSel][9];
    double N[9][9];
    hgeom::Norm nm;
};
struct PlaneScratch {
    float Fh[kThreads][9];
    uint32_t cnt[kThreads];
    uint32_t ok[kThreads];
    float4 tile[kTile];
};

struct Smem {
    uint32_t bits[kMaxWords];     // inlier bits of the last counted model
    uint32_t planar[kMaxWords];   // inlier bits of the current homography
    uint32_t prefix[kMaxWords + 1];
    uint32_t scan[kThreads];
    uint64_t key[kThreads];
    union {
        FitScratch fit;
        PlaneScratch plane;
    } u;
    uint32_t sel[kMaxSel];
    float4 sample[7];
    float models[3][9];
    float fitted[9];
    float curH[9];
    float fh[9];
    float bestF[9];
    float bestH[9];
    uint32_t acc;
    int flag;
    int nm;
    uint32_t nbPlanar;
};

__device__ inline uint32_t task_n(const DevTask& T) { return T.nDev ? min(*T.nDev, T.nHost) : T.nHost; }

__device__ inline uint64_t splitmix(uint64_t& s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline void copy9(float* dst, const float* src)
{
    if (threadIdx.x < 9) dst[threadIdx.x] = src[threadIdx.x];
}

// model kinds: 0 / 1 = F with the two-sided line / Sampson test, 2 = homography
template <int kKind>
__device__ inline bool is_inlier(const float m[9], float4 p, float t2)
{
    if constexpr (kKind == 2) return hgeom::inlier_h(m, p, t2);
    else return fgeom::inlier<kKind>(m, p, t2);
}

// inlier bits of model M (LDS) over the n points into dst; returns the count
template <int kKind>
__device__ uint32_t count_pass(Smem& s, const float4* pts, uint32_t n, const float* M, float t2, uint32_t* dst)
{
    float m[9];
#pragma unroll
    for (int q = 0; q < 9; q++) m[q] = M[q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) s.acc = 0;
    __syncthreads();
    uint32_t c = 0;
    // four rows of 256 points per trip, their loads issued before the tests
    constexpr uint32_t kU = 4;
    for (uint32_t base0 = 0; base0 < n; base0 += kU * kThreads) {
        float4 q[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint32_t i = base0 + u * kThreads + threadIdx.x;
            q[u] = i < n ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint32_t base = base0 + u * kThreads;
            if (base >= n) break;
            const uint32_t i = base + threadIdx.x;
            const bool in = i < n && is_inlier<kKind>(m, q[u], t2);
            const uint64_t b = __ballot(in);
            if (lane == 0) {
                const uint32_t w = (base >> 5) + 2 * wave;
                dst[w] = (uint32_t)b;
                dst[w + 1] = (uint32_t)(b >> 32);
                c += (uint32_t)__popcll(b);
            }
        }
    }
    if (lane == 0) atomicAdd(&s.acc, c);
    __syncthreads();
    const uint32_t r = s.acc;
    __syncthreads();
    return r;
}

// word w of the set selected from src: bits below n, optionally complemented
__device__ inline uint32_t word_of(const uint32_t* src, bool invert, uint32_t w, uint32_t n)
{
    uint32_t v = invert ? ~src[w] : src[w];
    const uint32_t lo = w * 32;
    if (lo + 32 > n) v &= (n > lo) ? ((1u << (n - lo)) - 1u) : 0u;
    return v;
}

// exclusive prefix of the set's word popcounts into s.prefix; returns the total
__device__ uint32_t prefix_scan(Smem& s, const uint32_t* src, bool invert, uint32_t n)
{
    const uint32_t nWords = (n + 31) / 32;
    constexpr uint32_t kPer = kMaxWords / kThreads;
    const uint32_t w0 = threadIdx.x * kPer;
    uint32_t local = 0;
    for (uint32_t k = 0; k < kPer; k++)
        if (w0 + k < nWords) local += (uint32_t)__popc(word_of(src, invert, w0 + k, n));
    s.scan[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const uint32_t v = threadIdx.x >= (uint32_t)off ? s.scan[threadIdx.x - off] : 0u;
        __syncthreads();
        s.scan[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s.scan[threadIdx.x] - local;
    for (uint32_t k = 0; k < kPer; k++)
        if (w0 + k < nWords) {
            s.prefix[w0 + k] = run;
            run += (uint32_t)__popc(word_of(src, invert, w0 + k, n));
        }
    if (threadIdx.x == kThreads - 1) s.prefix[nWords] = s.scan[kThreads - 1];
    __syncthreads();
    const uint32_t total = s.prefix[nWords];
    __syncthreads();
    return total;
}

// at most kMaxSel members of src (count of them = total), evenly spaced by
// rank: rank r is taken when r = floor(j total / kMaxSel) for some j, and
// lands in s.sel[j]; all of them when total <= kMaxSel. Returns the number.
__device__ uint32_t select_evenly(Smem& s, const uint32_t* src, uint32_t n, uint32_t total)
{
    prefix_scan(s, src, false, n);
    const uint32_t nWords = (n + 31) / 32;
    const bool all = total <= (uint32_t)kMaxSel;
    for (uint32_t w = threadIdx.x; w < nWords; w += kThreads) {
        uint32_t x = word_of(src, false, w, n);
        uint32_t r = s.prefix[w];
        while (x) {
            const uint32_t b = (uint32_t)__ffs(x) - 1u;
            x &= x - 1u;
            if (all) s.sel[r] = w * 32 + b;
            else {
                const uint32_t j = (r * (uint32_t)kMaxSel + total - 1u) / total;
                if (j < (uint32_t)kMaxSel && (j * total) / (uint32_t)kMaxSel == r) s.sel[j] = w * 32 + b;
            }
            r++;
        }
    }
    __syncthreads();
    return all ? total : (uint32_t)kMaxSel;
}

// index of the member of rank r (r < total) of the set scanned into s.prefix
__device__ inline uint32_t select_rank(const Smem& s, const uint32_t* src, bool invert, uint32_t n, uint32_t r)
{
    const uint32_t nWords = (n + 31) / 32;
    uint32_t lo = 0, hi = nWords;  // largest w with prefix[w] <= r
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s.prefix[mid] <= r) lo = mid;
        else hi = mid;
    }
    uint32_t x = word_of(src, invert, lo, n);
    for (uint32_t k = s.prefix[lo]; k < r; k++) x &= x - 1u;
    return min(lo * 32 + (uint32_t)__ffs(x) - 1u, n - 1u);  // the clamp never binds for r < total
}

// least-squares refit on the m >= 4 (H) / >= 8 (F) points in s.sel:
// s.fitted and s.flag
template <bool kH>
__device__ void fit_ls(Smem& s, const float4* pts, uint32_t m)
{
    FitScratch& f = s.u.fit;
    const uint32_t t = threadIdx.x;
    if (t < m) {
        const float4 p = pts[min(s.sel[t], kMaxPts - 1u)];
        f.pt[t][0] = p.x;
        f.pt[t][1] = p.y;
        f.pt[t][2] = p.z;
        f.pt[t][3] = p.w;
    }
    __syncthreads();
    const double dm = (double)m;
    if (t == 0) {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t j = 0; j < m; j++)
#pragma unroll
            for (int k = 0; k < 4; k++) a[k] = a[k] + f.pt[j][k];
#pragma unroll
        for (int k = 0; k < 4; k++) f.nm.c[k] = a[k] / dm;
    }
    __syncthreads();
    if (t < m) {
        const double dx0 = f.pt[t][0] - f.nm.c[0], dy0 = f.pt[t][1] - f.nm.c[1];
        const double dx1 = f.pt[t][2] - f.nm.c[2], dy1 = f.pt[t][3] - f.nm.c[3];
        f.d0[t] = sqrt(dx0 * dx0 + dy0 * dy0);
        f.d1[t] = sqrt(dx1 * dx1 + dy1 * dy1);
    }
    __syncthreads();
    if (t == 0) {
        double a0 = 0.0, a1 = 0.0;
        for (uint32_t j = 0; j < m; j++) {
            a0 = a0 + f.d0[j];
            a1 = a1 + f.d1[j];
        }
        const double n0 = a0 / dm, n1 = a1 / dm;
        s.flag = (n0 > 0.0) && (n1 > 0.0);
        f.nm.s0 = hgeom::kSqrt2 / n0;
        f.nm.s1 = hgeom::kSqrt2 / n1;
    }
    __syncthreads();
    if (!s.flag) return;
    if (t < m) {
        const double x0 = (f.pt[t][0] - f.nm.c[0]) * f.nm.s0, y0 = (f.pt[t][1] - f.nm.c[1]) * f.nm.s0;
        const double x1 = (f.pt[t][2] - f.nm.c[2]) * f.nm.s1, y1 = (f.pt[t][3] - f.nm.c[3]) * f.nm.s1;
        if constexpr (kH) {
            double* r1 = f.row[2 * t];
            double* r2 = f.row[2 * t + 1];
            r1[0] = -x0;
            r1[1] = -y0;
            r1[2] = -1.0;
            r1[3] = 0.0;
            r1[4] = 0.0;
            r1[5] = 0.0;
            r1[6] = x0 * x1;
            r1[7] = y0 * x1;
            r1[8] = x1;
            r2[0] = 0.0;
            r2[1] = 0.0;
            r2[2] = 0.0;
            r2[3] = -x0;
            r2[4] = -y0;
            r2[5] = -1.0;
            r2[6] = x0 * y1;
            r2[7] = y0 * y1;
            r2[8] = y1;
        } else {
            double* r = f.row[t];
            r[0] = x1 * x0;
            r[1] = y1 * x0;
            r[2] = x0;
            r[3] = x1 * y0;
            r[4] = y1 * y0;
            r[5] = y0;
            r[6] = x1;
            r[7] = y1;
            r[8] = 1.0;
        }
    }
    __syncthreads();
    if (t < 45) {
        // (a, b), a <= b, in row-major order of the upper triangle
        uint32_t a = 0, e = t;
        while (e >= 9 - a) {
            e -= 9 - a;
            a++;
        }
        const uint32_t b = a + e;
        const uint32_t rows = kH ? 2 * m : m;
        double acc = 0.0;
        for (uint32_t r = 0; r < rows; r++) acc = acc + f.row[r][a] * f.row[r][b];
        f.N[a][b] = acc;
        f.N[b][a] = acc;
    }
    __syncthreads();
    if (t == 0) {
        double v[9];
        bool ok = hgeom::null_vector9(f.N, v);
        if (ok) ok = kH ? hgeom::finish_h(v, f.nm, s.fitted) : hgeom::finish_f(v, f.nm, s.fitted);
        s.flag = ok;
    }
    __syncthreads();
}

// checkEpipolarityAndOptimise (degensac.cpp:65-78) on the F in M (LDS);
// M is replaced by the refit when that has more inliers. Returns the count.
template <int kMode>
__device__ uint32_t optimise_f(Smem& s, const float4* pts, uint32_t n, float* M, float t2)
{
    const uint32_t n0 = count_pass<kMode>(s, pts, n, M, t2, s.bits);
    if (n0 < 8) return n0;
    const uint32_t m = select_evenly(s, s.bits, n, n0);
    fit_ls<false>(s, pts, m);
    if (!s.flag) return n0;
    const uint32_t n1 = count_pass<kMode>(s, pts, n, s.fitted, t2, s.bits);
    if (n1 <= n0) return n0;
    copy9(M, s.fitted);
    __syncthreads();
    return n1;
}

// checkHomographyAndOptimise (degensac.cpp:87-100) on s.curH; its inlier
// bits end in s.planar. Returns the count.
__device__ uint32_t optimise_h(Smem& s, const float4* pts, uint32_t n, float t2)
{
    const uint32_t n0 = count_pass<2>(s, pts, n, s.curH, t2, s.planar);
    if (n0 < 4) return n0;
    const uint32_t m = select_evenly(s, s.planar, n, n0);
    fit_ls<true>(s, pts, m);
    if (!s.flag) return n0;
    const uint32_t n1 = count_pass<2>(s, pts, n, s.fitted, t2, s.bits);
    if (n1 <= n0) return n0;
    copy9(s.curH, s.fitted);
    for (uint32_t w = threadIdx.x; w < kMaxWords; w += kThreads) s.planar[w] = s.bits[w];
    __syncthreads();
    return n1;
}

__device__ uint64_t block_max(Smem& s, uint64_t v)
{
    s.key[threadIdx.x] = v;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s.key[threadIdx.x] = max(s.key[threadIdx.x], s.key[threadIdx.x + off]);
        __syncthreads();
    }
    const uint64_t r = s.key[0];
    __syncthreads();
    return r;
}

// 2-point sample sidx of candidate ordinal cand: ranks into the off-plane set
__device__ inline bool plane_sample(const Smem& s, const float4* pts, uint32_t n, uint32_t mOff, uint64_t seed,
                                   uint32_t cand, uint32_t sidx, float F[9])
{
    uint64_t st = seed ^ ((uint64_t)(cand * kMaxPlaneSamples + sidx) * 0xD1B54A32D192ED03ull);
    const uint32_t r0 = (uint32_t)(((splitmix(st) >> 32) * (uint64_t)mOff) >> 32);
    uint32_t r1 = (uint32_t)(((splitmix(st) >> 32) * (uint64_t)(mOff - 1u)) >> 32);
    if (r1 >= r0) r1++;
    const uint32_t ia = select_rank(s, s.planar, true, n, r0);
    const uint32_t ib = select_rank(s, s.planar, true, n, r1);
    return hgeom::f_from_h(s.curH, pts[ia], pts[ib], F);
}

// plane-and-parallax samples per workgroup: a candidate's nbPlaneSamples are
// split over ceil(nbPlaneSamples / kPlaneChunk) workgroups (16: the largest
// pairs' searches set the launch's length, 64 per workgroup was 3x slower)
constexpr uint32_t kPlaneChunk = 16;

// scores the plane-and-parallax samples [s0, s1) (s1 - s0 <= kThreads) over
// the points off s.curH (mOff >= 2 of them, inlier bits in s.planar, scanned
// into s.prefix): the samples' F in LDS, every wave takes all of them over
// its quarter of each point tile. Returns the best key count << 32 | ~sample
// (0 when no sample gave a usable F).
template <int kMode>
__device__ uint64_t plane_score(Smem& s, const float4* pts, uint32_t n, uint32_t mOff, const Params& p, uint32_t cand,
                                uint32_t s0, uint32_t s1)
{
    PlaneScratch& ps = s.u.plane;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t nHyp = s1 - s0;
    {
        const uint32_t sidx = s0 + threadIdx.x;
        float F[9];
        const bool ok = threadIdx.x < nHyp && plane_sample(s, pts, n, mOff, p.seed, cand, sidx, F);
#pragma unroll
        for (int q = 0; q < 9; q++) ps.Fh[threadIdx.x][q] = ok ? F[q] : 0.f;
        ps.ok[threadIdx.x] = ok;
        ps.cnt[threadIdx.x] = 0;
    }
    constexpr uint32_t kQuarter = kTile / 4;
    for (uint32_t base = 0; base < n; base += kTile) {
        const uint32_t tn = min((uint32_t)kTile, n - base);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < tn; i += kThreads) ps.tile[i] = pts[base + i];
        __syncthreads();
        const uint32_t q0 = wave * kQuarter, q1 = min(q0 + kQuarter, tn);
        for (uint32_t hh = 0; hh < nHyp; hh++) {
            if (!ps.ok[hh]) continue;
            float G[9];
#pragma unroll
            for (int q = 0; q < 9; q++) G[q] = ps.Fh[hh][q];
            uint32_t cnt = 0;
            for (uint32_t q = q0 + lane; q < q1; q += 64) {
                const uint32_t i = base + q;
                const bool off = !((s.planar[i >> 5] >> (i & 31)) & 1u);
                const bool in = off && fgeom::inlier<kMode>(G, ps.tile[q], p.tF2);
                cnt += (uint32_t)__popcll(__ballot(in));
            }
            if (lane == 0 && cnt) atomicAdd(&ps.cnt[hh], cnt);
        }
    }
    __syncthreads();
    const uint64_t key = ps.ok[threadIdx.x] ? (((uint64_t)ps.cnt[threadIdx.x] << 32) | (uint32_t)~(s0 + threadIdx.x)) : 0ull;
    return block_max(s, key);
}

// the nCand best samples (count, then lowest index; count > 0) of each pair
template <class Src>
__global__ __launch_bounds__(kThreads) void select_kernel(const Src batch, const Params p)
{
    __shared__ uint64_t sKey[kThreads];
    const DevTask& T = batch[blockIdx.x];
    const uint32_t n = task_n(T);
    uint64_t prevKey = ~0ull;
    for (uint32_t c = 0; c < p.nCand; c++) {
        uint64_t key = 0;
        if (n >= 7)
            for (uint32_t h = threadIdx.x; h < T.H; h += kThreads) {
                const uint32_t cnt = T.counts[h];
                const uint64_t k = ((uint64_t)cnt << 32) | (uint32_t)~h;
                if (cnt > 0 && k < prevKey && k > key) key = k;
            }
        sKey[threadIdx.x] = key;
        __syncthreads();
        for (int off = kThreads / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) sKey[threadIdx.x] = max(sKey[threadIdx.x], sKey[threadIdx.x + off]);
            __syncthreads();
        }
        key = sKey[0];
        __syncthreads();
        if (threadIdx.x == 0) T.cand[c].h = key == 0 ? ~0u : ~(uint32_t)key;
        prevKey = key == 0 ? 0ull : key;  // nothing is below 0: the rest stay empty
    }
}

// the per-new-best body of degensac.cpp:126-194 for candidate blockIdx.x of
// pair blockIdx.y, as if it beat the best so far (through the optimised H);
// its record in T.cand
template <int kMode, class Src>
__global__ __launch_bounds__(kThreads) void candidate_kernel(const Src batch, const Params p)
{
    __shared__ Smem s;
    const DevTask& T = batch[blockIdx.y];
    const uint32_t c = blockIdx.x;
    CandRec& R = T.cand[c];
    const uint32_t h = R.h;
    if (h == ~0u) return;
    const uint32_t n = task_n(T);
    const float4* pts = T.pts;
    bool inRange = true;
    for (int i = 0; i < 7; i++) inRange = inRange && T.idx[(size_t)h * 7 + i] < n;
    if (threadIdx.x < 7) s.sample[threadIdx.x] = pts[inRange ? T.idx[(size_t)h * 7 + threadIdx.x] : 0];
    __syncthreads();
    if (threadIdx.x == 0) {
        float4 q[7];
        for (int i = 0; i < 7; i++) q[i] = s.sample[i];
        float Fs[3][9];
        const int nm = inRange ? fgeom::fit7(q, Fs) : 0;
        for (int m = 0; m < nm; m++)
            for (int k = 0; k < 9; k++) s.models[m][k] = Fs[m][k];
        s.nm = nm;
    }
    __syncthreads();
    const int nm = s.nm;
    uint32_t cntF = 0;
    int bm = -1;
    for (int m = 0; m < nm; m++) {
        const uint32_t cm = optimise_f<kMode>(s, pts, n, s.models[m], p.tF2);
        if (cm > cntF) {
            cntF = cm;
            bm = m;
        }
    }
    if (threadIdx.x == 0) {
        R.cntF = cntF;
        R.planar = 0;
        R.cntH = 0;
        R.plane = 0;
        R.cntFh = 0;
        R.planeKey = 0;
    }
    if (bm < 0) return;
    copy9(s.bestF, s.models[bm]);
    copy9(R.F, s.models[bm]);
    __syncthreads();
    // degeneracy test of the sample under its best F
    if (threadIdx.x == 0) {
        uint32_t nbPlanar = 3;
        for (int i = 0; i < 5; i++) {
            const float4 tri[3] = {s.sample[kTri[i][0]], s.sample[kTri[i][1]], s.sample[kTri[i][2]]};
            float Hc[9];
            if (!hgeom::h_from_f(s.bestF, tri, Hc)) continue;
            uint32_t cc = 3;
            for (int k = 0; k < 4; k++) cc += hgeom::inlier_h(Hc, s.sample[kRem[i][k]], p.tH2) ? 1u : 0u;
            if (cc > nbPlanar) {
                nbPlanar = cc;
                for (int q = 0; q < 9; q++) s.curH[q] = Hc[q];
            }
        }
        s.nbPlanar = nbPlanar;
    }
    __syncthreads();
    if (s.nbPlanar < 5) return;
    const uint32_t cntH = optimise_h(s, pts, n, p.tH2);
    if (threadIdx.x == 0) {
        R.planar = 1;
        R.cntH = cntH;
    }
    copy9(R.H, s.curH);
}

// whether candidate c can become the best at all: the best count before it
// is at least every earlier candidate's count, so it needs a count above all
// of them (a necessary condition; the replay decides)
__device__ inline bool may_lead(const DevTask& T, uint32_t c)
{
    uint32_t before = 0;
    for (uint32_t k = 0; k < c; k++) before = max(before, T.cand[k].cntF);
    return T.cand[c].cntF > before;
}

// the off-plane set of candidate c's optimised H again: inlier bits into
// s.planar (s.curH = the stored H) and their complement scanned into
// s.prefix; returns its size
__device__ uint32_t off_plane(Smem& s, const CandRec& R, const float4* pts, uint32_t n, const Params& p)
{
    if (threadIdx.x < 9) s.curH[threadIdx.x] = R.H[threadIdx.x];
    __syncthreads();
    count_pass<2>(s, pts, n, s.curH, p.tH2, s.planar);
    return prefix_scan(s, s.planar, true, n);
}

// the plane-and-parallax search (degensac.cpp:195-228) of a planar
// candidate, its samples split over workgroups: workgroup x of pair y scores
// samples [kPlaneChunk g, kPlaneChunk (g + 1)) of candidate x / nChunks
// (g = x % nChunks) and merges its best key into the record
template <int kMode, class Src>
__global__ __launch_bounds__(kThreads) void plane_kernel(const Src batch, const Params p, uint32_t nChunks)
{
    __shared__ Smem s;
    const DevTask& T = batch[blockIdx.y];
    const uint32_t c = blockIdx.x / nChunks, g = blockIdx.x % nChunks;
    CandRec& R = T.cand[c];
    if (R.h == ~0u || !R.planar || !may_lead(T, c)) return;
    const uint32_t n = task_n(T);
    const uint32_t mOff = off_plane(s, R, T.pts, n, p);
    if (mOff < 2) return;
    const uint32_t s0 = g * kPlaneChunk, s1 = min(s0 + kPlaneChunk, p.nPlane);
    const uint64_t key = plane_score<kMode>(s, T.pts, n, mOff, p, c, s0, s1);
    if (threadIdx.x == 0 && key != 0)
        atomicMax(reinterpret_cast<unsigned long long*>(&R.planeKey), (unsigned long long)key);
}

// the best plane-and-parallax F of a candidate, optimised on all points
template <int kMode, class Src>
__global__ __launch_bounds__(kThreads) void plane_finish_kernel(const Src batch, const Params p)
{
    __shared__ Smem s;
    const DevTask& T = batch[blockIdx.y];
    const uint32_t c = blockIdx.x;
    CandRec& R = T.cand[c];
    if (R.h == ~0u || !R.planar || !may_lead(T, c)) return;
    const uint64_t key = R.planeKey;
    if (key == 0) return;
    const uint32_t n = task_n(T);
    const float4* pts = T.pts;
    const uint32_t mOff = off_plane(s, R, pts, n, p);
    if (threadIdx.x == 0) {
        float F[9];
        plane_sample(s, pts, n, mOff, p.seed, c, ~(uint32_t)key, F);
#pragma unroll
        for (int q = 0; q < 9; q++) s.fh[q] = F[q];
    }
    __syncthreads();
    const uint32_t cFh = optimise_f<kMode>(s, pts, n, s.fh, p.tF2);
    if (threadIdx.x == 0) {
        R.plane = 1;
        R.cntFh = cFh;
    }
    copy9(R.Fh, s.fh);
}

// the reference's sequence over the records (degensac.cpp:126-231): a
// candidate runs its body only when its count beats the best so far, and its
// H and plane-and-parallax F replace the best ones when they have more
// inliers; then the results and both inlier masks
template <int kMode, class Src>
__global__ __launch_bounds__(kThreads) void replay_kernel(const Src batch, const Params p)
{
    __shared__ float sF[9], sH[9];
    __shared__ uint32_t sRes[4];
    const DevTask& T = batch[blockIdx.x];
    const uint32_t n = task_n(T);
    const float4* pts = T.pts;
    if (threadIdx.x == 0) {
        uint32_t bestCntF = 0, bestCntH = 0, source = 0, bestHyp = ~0u;
        const float* bF = nullptr;
        const float* bH = nullptr;
        for (uint32_t c = 0; n >= 7 && c < p.nCand; c++) {
            const CandRec& R = T.cand[c];
            if (R.h == ~0u) break;
            if (R.cntF <= bestCntF) continue;
            bestCntF = R.cntF;
            source = 0;
            bestHyp = R.h;
            bF = R.F;
            if (!R.planar) continue;
            if (R.cntH > bestCntH) {
                bestCntH = R.cntH;
                bH = R.H;
            }
            if (!R.plane) continue;
            if (R.cntFh > bestCntF) {
                bestCntF = R.cntFh;
                source = 1;
                bF = R.Fh;
            }
        }
        for (int q = 0; q < 9; q++) {
            sF[q] = bF ? bF[q] : 0.f;
            sH[q] = bH ? bH[q] : 0.f;
        }
        sRes[0] = bestCntF;
        sRes[1] = bestCntH;
        sRes[2] = source;
        sRes[3] = bestHyp;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        T.F[threadIdx.x] = sF[threadIdx.x];
        T.Hm[threadIdx.x] = sH[threadIdx.x];
    }
    if (threadIdx.x < 4) T.result[threadIdx.x] = sRes[threadIdx.x];
    const uint32_t bestCntF = sRes[0], bestCntH = sRes[1];
    float F[9], Hh[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        F[q] = sF[q];
        Hh[q] = sH[q];
    }
    for (uint32_t i = threadIdx.x; i < T.nHost; i += kThreads) {
        const bool live = i < n;
        if (T.maskF) T.maskF[i] = (live && bestCntF > 0) ? (uint8_t)fgeom::inlier<kMode>(F, pts[i], p.tF2) : 0;
        if (T.maskH) T.maskH[i] = (live && bestCntH > 0) ? (uint8_t)hgeom::inlier_h(Hh, pts[i], p.tH2) : 0;
    }
}

template <int kMode, class Src>
static int launch_all(const Src& src, uint32_t nTasks, const Params& p, hipStream_t stream)
{
    select_kernel<<<nTasks, kThreads, 0, stream>>>(src, p);
    RSFM_CHECK(check_launch("degensac select"));
    candidate_kernel<kMode><<<dim3(p.nCand, nTasks), kThreads, 0, stream>>>(src, p);
    RSFM_CHECK(check_launch("degensac candidates"));
    const uint32_t nChunks = (p.nPlane + kPlaneChunk - 1) / kPlaneChunk;
    plane_kernel<kMode><<<dim3(p.nCand * nChunks, nTasks), kThreads, 0, stream>>>(src, p, nChunks);
    RSFM_CHECK(check_launch("degensac plane search"));
    plane_finish_kernel<kMode><<<dim3(p.nCand, nTasks), kThreads, 0, stream>>>(src, p);
    RSFM_CHECK(check_launch("degensac plane finish"));
    replay_kernel<kMode><<<nTasks, kThreads, 0, stream>>>(src, p);
    return check_launch("degensac replay");
}

}  // namespace degensac
}  // namespace rsfm_amd

using namespace rsfm_amd;

namespace rsfm_amd {
namespace degensac {

int make_params(float thresholdF, float thresholdH, int mode, uint32_t nbCandidates, uint32_t nbPlaneSamples,
                uint64_t seed, Params& out)
{
    RSFM_RETURN_IF(mode != 0 && mode != 1, RSFM_ERR_INVALID, "degensac: mode must be 0 (two-sided line) or 1 (Sampson)");
    RSFM_RETURN_IF(!(thresholdF > 0.f) || !(thresholdH > 0.f), RSFM_ERR_INVALID, "degensac: thresholds must be > 0");
    RSFM_RETURN_IF(nbCandidates < 1 || nbCandidates > kMaxCand, RSFM_ERR_INVALID,
                   "degensac: nbCandidates must be in [1, 64]");
    RSFM_RETURN_IF(nbPlaneSamples < 1 || nbPlaneSamples > kMaxPlaneSamples, RSFM_ERR_INVALID,
                   "degensac: nbPlaneSamples must be in [1, 4096]");
    out = Params{thresholdF * thresholdF, thresholdH * thresholdH, nbCandidates, nbPlaneSamples, seed};
    return RSFM_OK;
}

int make_task(const rsfm_degensac_task& t, DevTask& out)
{
    RSFM_RETURN_IF(t.nbPts > kMaxPts, RSFM_ERR_UNSUPPORTED, "degensac: at most 65536 correspondences per pair");
    RSFM_RETURN_IF(t.nbPts > 0 && t.pts == nullptr, RSFM_ERR_INVALID, "degensac: pts is null");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16, RSFM_ERR_INVALID, "degensac: pts must be 16-byte aligned");
    RSFM_RETURN_IF(t.nbHyp > 0 && (t.sampleIdx == nullptr || t.counts == nullptr), RSFM_ERR_INVALID,
                   "degensac: sampleIdx/counts must be set when nbHyp > 0");
    RSFM_RETURN_IF(t.F == nullptr || t.H == nullptr || t.result == nullptr, RSFM_ERR_INVALID,
                   "degensac: F, H and result are required");
    out = DevTask{reinterpret_cast<const float4*>(t.pts), t.nbPtsDev, t.nbPts, t.sampleIdx, t.counts, t.nbHyp, t.F, t.H,
                  t.result, t.inlierMaskF, t.inlierMaskH};
    return RSFM_OK;
}

int launch_degensac_table(const DevTask* tasks, uint32_t nTasks, const Params& p, int mode, hipStream_t stream)
{
    if (nTasks == 0) return RSFM_OK;
    const InTable<DevTask> src{tasks};
    return mode == 0 ? launch_all<0>(src, nTasks, p, stream) : launch_all<1>(src, nTasks, p, stream);
}

}  // namespace degensac
}  // namespace rsfm_amd

extern "C" int rsfm_degensac_refine(const rsfm_degensac_task* tasks, size_t nbTasks, float thresholdF,
                                    float thresholdH, int mode, uint32_t nbCandidates, uint32_t nbPlaneSamples,
                                    uint64_t seed, hipStream_t stream)
{
    using namespace degensac;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "degensac: tasks is null");
    Params prm;
    RSFM_CHECK(make_params(thresholdF, thresholdH, mode, nbCandidates, nbPlaneSamples, seed, prm));
    // the candidate records come from the stream-ordered allocator (the
    // reference interface has no workspace)
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        Batch batch{};
        for (size_t i = 0; i < n; i++) RSFM_CHECK(make_task(tasks[b + i], batch.t[i]));
        void* ws = nullptr;
        RSFM_CHECK(check_hip(hipMallocAsync(&ws, n * prm.nCand * sizeof(CandRec), stream), "degensac: candidate records"));
        for (size_t i = 0; i < n; i++) batch.t[i].cand = static_cast<CandRec*>(ws) + i * prm.nCand;
        int rc = mode == 0 ? launch_all<0>(batch, (uint32_t)n, prm, stream) : launch_all<1>(batch, (uint32_t)n, prm, stream);
        const int rf = check_hip(hipFreeAsync(ws, stream), "degensac: candidate records release");
        if (rc == RSFM_OK) rc = rf;
        RSFM_CHECK(rc);
    }
    return RSFM_OK;
}
