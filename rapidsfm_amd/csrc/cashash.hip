// cashash.hip - cascade hashing on gfx950.
//
// Reference (rsfm::cashash, casHash/cascadeHash.h:32-107):
//  accumuateDesc     casHash/accumulateDesc.cu:33-131   acc[k] += byte_k >> 1 over all descriptors
//  computeBias       casHash/accumulateDesc.cu:136-182  bias[c] = round(sum_k w[c][k] * -acc[k]/n) + biasBase[c]
//  computeHash*      casHash/computeHash.cu:437-520     bit c = bias[c] + sum_k w[c][k] (byte_k >> 1) >= 0
//  buildHistogram    casHash/buildBuckets.cu:29-114
//  computeBucketBound casHash/buildBuckets.cu:116-172
//  buildBuckets      casHash/buildBuckets.cu:178-269    bucket CSR (order inside a bucket unspecified)
//  findBestMatch     casHash/findBestMatch.cu:570-676   exact L2^2 over the union of same-bucket candidates
//
// GPU design (own):
//  * the hash projection is an [n x 128] x [128 x nch] int8 GEMM on
//    v_mfma_i32_32x32x32_i8; after adding the bias, one wave ballot per
//    accumulator register yields 32 channel bits for two descriptors at once,
//    already in the code-byte / remap-word bit order.
//  * histogram, bounds and buckets use LDS atomics and one global reservation
//    per (scheme, bucket) per workgroup.
//  * matching cuts each (scheme, bucket) candidate set into 64 x 64 blocks
//    on v_mfma_i32_16x16x64_i8 (against 128 x 128 tiles in the reference),
//    balanced over all waves of the launch; row and column minima are folded
//    in registers with 32-bit (distance, position) keys inside a block and
//    merged across blocks with 64-bit atomicMin on (distance, global index)
//    keys: once per query row and run of blocks, once per train column and
//    block.
#include "common.hpp"
#include <algorithm>
#include <vector>

namespace rsfm_amd {
namespace cashash {

constexpr uint32_t kSchemes = 8;
constexpr uint32_t kBuckets = 256;

// ---------------------------------------------------------------------------
// accumulate (byte >> 1) per dimension into u64[128]
__global__ void zero_u64_kernel(uint64_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

constexpr int kAccThreads = 256;
constexpr int kAccDescPerBlock = 64;  // 128 workgroups for 8K descriptors

__global__ __launch_bounds__(kAccThreads) void accumulate_kernel(uint64_t* __restrict__ acc, uint64_t* __restrict__ sqrAcc,
                                                                 const uint8_t* __restrict__ desc, uint32_t n)
{
    // thread -> (16-byte column group tid % 8, descriptor subset tid / 8)
    __shared__ uint32_t sAcc[128];
    __shared__ uint32_t sSqr[128];
    if (threadIdx.x < 128) {
        sAcc[threadIdx.x] = 0;
        sSqr[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t part = threadIdx.x & 7, sub = threadIdx.x >> 3;
    uint32_t a[16], q[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        a[k] = 0;
        q[k] = 0;
    }
    const uint32_t d0 = blockIdx.x * kAccDescPerBlock;
    for (uint32_t d = d0 + sub; d < min(n, d0 + kAccDescPerBlock); d += kAccThreads / 8) {
        const uint4 x = *reinterpret_cast<const uint4*>(desc + (size_t)d * 128 + part * 16);
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint32_t v = ((w[j] >> (8 * b)) & 0xFF) >> 1;
                a[j * 4 + b] += v;
                q[j * 4 + b] += v * v;
            }
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
        atomicAdd(&sAcc[part * 16 + k], a[k]);
        if (sqrAcc) atomicAdd(&sSqr[part * 16 + k], q[k]);
    }
    __syncthreads();
    if (threadIdx.x < 128) {
        atomicAdd(reinterpret_cast<unsigned long long*>(acc + threadIdx.x), (unsigned long long)sAcc[threadIdx.x]);
        if (sqrAcc)
            atomicAdd(reinterpret_cast<unsigned long long*>(sqrAcc + threadIdx.x), (unsigned long long)sSqr[threadIdx.x]);
    }
}

// ---------------------------------------------------------------------------
// bias: FP32, 8 lanes per channel each summing 16 products in k order with
// fused multiply-add, then an xor tree over the 8 lanes (4, 2, 1)
__global__ void bias_kernel(int32_t* __restrict__ bias, const int8_t* __restrict__ w, const int32_t* __restrict__ biasBase,
                            const uint64_t* __restrict__ acc, uint64_t n, uint32_t nch)
{
    __shared__ float sNeg[128];
    if (threadIdx.x < 128) sNeg[threadIdx.x] = -(float)acc[threadIdx.x] / (float)n;
    __syncthreads();
    const uint32_t c = blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3), L = threadIdx.x & 7;
    if (c >= nch) return;  // whole groups of 8 lanes leave together
    float v = 0.f;
#pragma unroll
    for (int k = 0; k < 16; k++) v = __builtin_fmaf(sNeg[L * 16 + k], (float)w[c * 128 + L * 16 + k], v);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 1);
    if (L == 0) bias[c] = (int32_t)roundf(v) + biasBase[c];
}

// ---------------------------------------------------------------------------
// hash: one wave per 32 descriptors, all channel blocks of 32
constexpr int kHashWaves = 4;

template <int kChBlocks>
__global__ __launch_bounds__(kHashWaves * 64) void hash_kernel(const int8_t* __restrict__ w, const int32_t* __restrict__ bias,
                                                                const uint8_t* __restrict__ desc, uint32_t n,
                                                                uint32_t* __restrict__ codes, uint32_t* __restrict__ remap,
                                                                uint32_t* __restrict__ hist)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l32 = lane & 31, hi = lane >> 5;
    const uint32_t rowBase = (blockIdx.x * kHashWaves + wave) * 32;
    if (rowBase >= n) return;
    const uint32_t row = rowBase + l32;
    const bool valid = row < n;
    v4i aop[4];
#pragma unroll
    for (int kc = 0; kc < 4; kc++) {
        v4i x = {0, 0, 0, 0};
        if (valid) x = *reinterpret_cast<const v4i*>(desc + (size_t)row * 128 + kc * 32 + hi * 16);
#pragma unroll
        for (int q = 0; q < 4; q++) x[q] = (int)(((uint32_t)x[q] >> 1) & 0x7F7F7F7Fu);
        aop[kc] = x;
    }
    uint32_t myWord[kChBlocks];
#pragma unroll
    for (int cb = 0; cb < kChBlocks; cb++) {
        const int ch = cb * 32 + l32;
        v16i acc = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            const v4i bop = *reinterpret_cast<const v4i*>(w + (size_t)ch * 128 + kc * 32 + hi * 16);
            acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[kc], bop, acc, 0, 0, 0);
        }
        const int32_t bc = bias[ch];
        uint32_t mine = 0;
#pragma unroll
        for (int v = 0; v < 16; v++) {
            // lanes 0..31 hold row r0 = 8(v/4)+(v%4), lanes 32..63 row r0 + 4;
            // lane%32 is the channel inside this block
            const uint64_t b = __ballot(acc[v] + bc >= 0);
            const int r0 = 8 * (v >> 2) + (v & 3);
            if (l32 == r0) mine = (uint32_t)b;
            if (l32 == r0 + 4) mine = (uint32_t)(b >> 32);
        }
        myWord[cb] = mine;
    }
    if (valid && hi == 0) {
        codes[(size_t)row * 2 + 0] = myWord[0];
        codes[(size_t)row * 2 + 1] = myWord[1];
        if constexpr (kChBlocks > 2) {
            if (remap) {
#pragma unroll
                for (int r = 0; r < 4; r++) remap[(size_t)row * 4 + r] = myWord[2 + r];
            }
        }
        if (hist) {
#pragma unroll
            for (int s = 0; s < 8; s++) atomicAdd(&hist[s * 256 + ((myWord[s >> 2] >> (8 * (s & 3))) & 0xFF)], 1u);
        }
    }
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
constexpr int kHistThreads = 256;
constexpr int kHistPerBlock = 512;

__global__ __launch_bounds__(kHistThreads) void histogram_kernel(uint32_t* __restrict__ hist, const uint8_t* __restrict__ codes,
                                                                 uint32_t n)
{
    __shared__ uint32_t sH[kSchemes * kBuckets];
    for (uint32_t i = threadIdx.x; i < kSchemes * kBuckets; i += kHistThreads) sH[i] = 0;
    __syncthreads();
    const uint32_t d0 = blockIdx.x * kHistPerBlock;
    for (uint32_t d = d0 + threadIdx.x; d < min(n, d0 + kHistPerBlock); d += kHistThreads) {
        const uint2 c = *reinterpret_cast<const uint2*>(codes + (size_t)d * 8);
#pragma unroll
        for (int s = 0; s < 8; s++) atomicAdd(&sH[s * 256 + (((s < 4 ? c.x : c.y) >> (8 * (s & 3))) & 0xFF)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kSchemes * kBuckets; i += kHistThreads)
        if (sH[i]) atomicAdd(&hist[i], sH[i]);
}

// exclusive (lower) and inclusive (upper) prefix per scheme, largest bucket:
// wave s scans scheme s (four buckets per lane, then a wave prefix over the
// lane totals), so the eight scans run side by side
constexpr int kBoundsThreads = 64 * kSchemes;

__global__ __launch_bounds__(kBoundsThreads) void bounds_kernel(uint32_t* __restrict__ lower, uint32_t* __restrict__ upper,
                                                                uint32_t* __restrict__ maxBucket,
                                                                const uint32_t* __restrict__ hist)
{
    __shared__ uint32_t sMax[kSchemes];
    const uint32_t s = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t* hs = hist + s * 256 + lane * 4;
    const uint4 h = make_uint4(hs[0], hs[1], hs[2], hs[3]);
    const uint32_t p0 = h.x, p1 = p0 + h.y, p2 = p1 + h.z, p3 = p2 + h.w;
    uint32_t scan = p3;  // inclusive prefix of the lane totals
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(scan, off);
        if (lane >= (uint32_t)off) scan += v;
    }
    const uint32_t base = scan - p3;
    uint32_t* up = upper + s * 256 + lane * 4;
    uint32_t* lo = lower + s * 256 + lane * 4;
    up[0] = base + p0;
    up[1] = base + p1;
    up[2] = base + p2;
    up[3] = base + p3;
    lo[0] = base;
    lo[1] = base + p0;
    lo[2] = base + p1;
    lo[3] = base + p2;
    uint32_t mx = max(max(h.x, h.y), max(h.z, h.w));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off));
    if (lane == 0) sMax[s] = mx;
    __syncthreads();
    if (threadIdx.x == 0 && maxBucket) {
        uint32_t m = 0;
        for (uint32_t k = 0; k < kSchemes; k++) m = max(m, sMax[k]);
        *maxBucket = m;
    }
}

// two-pass scatter: block histogram, one global slot reservation per
// (scheme, bucket), then LDS slot counters
constexpr int kBuildThreads = 256;
constexpr int kBuildPerBlock = 256;

__global__ __launch_bounds__(kBuildThreads) void buckets_kernel(const uint8_t* __restrict__ codes, uint32_t n,
                                                               uint32_t* __restrict__ alloc, uint32_t* __restrict__ buckets)
{
    __shared__ uint32_t sH[kSchemes * kBuckets];
    for (uint32_t i = threadIdx.x; i < kSchemes * kBuckets; i += kBuildThreads) sH[i] = 0;
    __syncthreads();
    const uint32_t d0 = blockIdx.x * kBuildPerBlock;
    const uint32_t d1 = min(n, d0 + kBuildPerBlock);
    for (uint32_t d = d0 + threadIdx.x; d < d1; d += kBuildThreads) {
        const uint2 c = *reinterpret_cast<const uint2*>(codes + (size_t)d * 8);
#pragma unroll
        for (int s = 0; s < 8; s++) atomicAdd(&sH[s * 256 + (((s < 4 ? c.x : c.y) >> (8 * (s & 3))) & 0xFF)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kSchemes * kBuckets; i += kBuildThreads)
        sH[i] = sH[i] ? atomicAdd(&alloc[i], sH[i]) : 0u;
    __syncthreads();
    for (uint32_t d = d0 + threadIdx.x; d < d1; d += kBuildThreads) {
        const uint2 c = *reinterpret_cast<const uint2*>(codes + (size_t)d * 8);
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const uint32_t b = ((s < 4 ? c.x : c.y) >> (8 * (s & 3))) & 0xFF;
            const uint32_t slot = atomicAdd(&sH[s * 256 + b], 1u);
            buckets[(size_t)s * n + slot] = d;
        }
    }
}

// ---------------------------------------------------------------------------
// matching
struct DevImage {
    const uint8_t* desc;
    const uint32_t* ends;     // [8][256] inclusive prefix
    const uint32_t* buckets;  // [8][nDesc]
    uint32_t n;
};
constexpr int kMatchMaxTasks = 16;
struct MatchBatch {
    DevImage q[kMatchMaxTasks];
    uint64_t* resQ[kMatchMaxTasks];
    uint64_t* resT[kMatchMaxTasks];
    DevImage t;
};

// Work split. Bucket sizes are far from uniform (with the reference's
// projection construction the largest of 256 buckets can hold half an
// image), so a task is cut into blocks of at most 64 query x 64 train
// members of one bucket, ordered (scheme, bucket, query chunk, train chunk),
// and every wave of the launch takes an even slice of its task's blocks. A
// wave keeps its query chunk (operands in registers, row minima as
// (distance, index) keys) while its consecutive blocks share it and merges
// each row into the result once per run; each block's train-side minima are
// merged once per train member. Per block, on v_mfma_i32_16x16x64_i8 (lane
// l: column class l % 16, row group q = l / 16; products of rows
// rb * 16 + 4 q + j and columns cb * 16 + l % 16):
//  * both chunks' member indices are sorted in the wave (bitonic over the 64
//    lanes), so a position inside a chunk orders like the global index and
//    32-bit keys (distance << 6 | position) break ties as the reference does;
//  * operands are gathered straight into registers: u = 127 - a for query
//    rows, w = b - 128 for train columns (zero for positions past the
//    chunk), L2^2 = NA' + NB' + 2 u.w with NA' = |a - 128|^2 and
//    NB' = |b - 127|^2 - 128;
//  * row key (acc << 7) + (NB' << 6 | column) and column key
//    (acc << 7) + (NA' << 6 | row), one v_lshl_add each; a row's minimum over
//    the lane's 4 columns goes through a per-wave LDS transpose, after which
//    lane L takes row L's 16 class minima and adds NA'; a column's minimum
//    over the lane's 16 rows, two lane swaps that leave lane L with column L,
//    then NB' added.
constexpr int kMatchWaves = 4;
constexpr int kRowStride = 17;
constexpr int32_t kPastChunk = 0x00FC0000;   // norm of a position past the chunk: key 0x3F000000 | position
constexpr int32_t kFoundLimit = 1 << 29;     // real keys: distance << 6 | position, distance < 2^23

__device__ inline uint32_t wave_sort_u32(uint32_t v, int lane)
{
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)v, j);
            const bool keepMin = ((lane & j) == 0) == ((lane & k) == 0);
            v = keepMin ? min(v, o) : max(v, o);
        }
    return v;
}

// a value the whole wave holds (an LDS word read at a wave-uniform address)
__device__ inline uint32_t uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }

// (a << 7) + b in plain code: a is an MFMA accumulator, and only compiled
// code gets the wait states a VALU read of an MFMA result needs (inline asm
// reading it could see a stale register)
__device__ inline int32_t shl7_add(int32_t a, int32_t b) { return (int32_t)(((uint32_t)a << 7) + (uint32_t)b); }

// __launch_bounds__(256, 3): at least three waves per SIMD (at most 168 VGPRs)
__global__ __launch_bounds__(kMatchWaves * 64, 3) void match_kernel(const MatchBatch batch, uint32_t nSchemes)
{
    const uint32_t task = blockIdx.z;
    const DevImage& Q = batch.q[task];
    const DevImage& T = batch.t;
    uint64_t* resQ = batch.resQ[task];
    uint64_t* resT = batch.resT[task];
    if (Q.n == 0 || T.n == 0) return;
    // flat bucket index fb = scheme * 256 + bucket
    __shared__ uint32_t sPre[kSchemes * kBuckets + 1];  // first block of each bucket
    __shared__ uint32_t sQE[kSchemes * kBuckets], sTE[kSchemes * kBuckets];  // inclusive bucket ends
    __shared__ uint32_t sWave[kSchemes][kMatchWaves];
    __shared__ int32_t sRow[kMatchWaves][64 * kRowStride];  // per wave: [row][column class]
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c16 = lane & 15, q = lane >> 4;
    static_assert(kMatchWaves * 64 == kBuckets, "one thread per bucket");
    {
        // blocks per bucket, ceil(q / 64) * ceil(t / 64), and their exclusive
        // prefix over all (scheme, bucket) in order; every load issued first
        uint32_t qe[kSchemes], qb[kSchemes], te[kSchemes], tb[kSchemes];
#pragma unroll
        for (uint32_t s = 0; s < kSchemes; s++) {
            if (s >= nSchemes) continue;
            const uint32_t o = s * kBuckets + (uint32_t)tid;
            qe[s] = Q.ends[o];
            te[s] = T.ends[o];
            qb[s] = tid ? Q.ends[o - 1] : 0u;
            tb[s] = tid ? T.ends[o - 1] : 0u;
        }
        uint32_t cnt[kSchemes], inc[kSchemes];
#pragma unroll
        for (uint32_t s = 0; s < kSchemes; s++) {
            if (s >= nSchemes) continue;
            sQE[s * kBuckets + tid] = qe[s];
            sTE[s * kBuckets + tid] = te[s];
            cnt[s] = ((qe[s] - qb[s] + 63) >> 6) * ((te[s] - tb[s] + 63) >> 6);
            inc[s] = cnt[s];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t v = (uint32_t)__shfl_up((int)inc[s], off);
                if (lane >= off) inc[s] += v;
            }
            if (lane == 63) sWave[s][wave] = inc[s];
        }
        __syncthreads();
        uint32_t base = 0;
#pragma unroll
        for (uint32_t s = 0; s < kSchemes; s++) {
            if (s >= nSchemes) continue;
            uint32_t before = 0, all = 0;
#pragma unroll
            for (int w = 0; w < kMatchWaves; w++) {
                before += w < wave ? sWave[s][w] : 0u;
                all += sWave[s][w];
            }
            sPre[s * kBuckets + tid] = base + before + inc[s] - cnt[s];
            base += all;
        }
        if (tid == 0) sPre[nSchemes * kBuckets] = base;
    }
    __syncthreads();
    const uint32_t nFb = nSchemes * kBuckets;
    const uint32_t total = uniform(sPre[nFb]);
    const uint32_t gw = blockIdx.x * kMatchWaves + (uint32_t)wave, nW = gridDim.x * kMatchWaves;
    uint32_t k = (uint32_t)(((uint64_t)gw * total) / nW);
    const uint32_t kEnd = (uint32_t)(((uint64_t)(gw + 1) * total) / nW);
    if (k >= kEnd) return;  // no barrier follows
    // bucket of block k: the largest fb with sPre[fb] <= k
    uint32_t fb = 0;
    {
        uint32_t hiB = nFb;
        while (hiB - fb > 1) {
            const uint32_t mid = (fb + hiB) >> 1;
            if (uniform(sPre[mid]) <= k) fb = mid;
            else hiB = mid;
        }
    }
    uint32_t qBeg, qCnt, tBeg, tCnt;
    auto bucket = [&]() {
        const bool first = (fb & (kBuckets - 1)) == 0;
        qBeg = first ? 0u : uniform(sQE[fb - 1]);
        tBeg = first ? 0u : uniform(sTE[fb - 1]);
        qCnt = uniform(sQE[fb]) - qBeg;
        tCnt = uniform(sTE[fb]) - tBeg;
    };
    bucket();
    uint32_t nT = (tCnt + 63) >> 6, nQ = (qCnt + 63) >> 6;
    const uint32_t kIn = k - uniform(sPre[fb]);
    uint32_t qt = kIn / nT, tt = kIn - qt * nT;
    const bool doCols = resT != nullptr;
    const uint8_t* qDesc = Q.desc;
    const uint8_t* tDesc = T.desc;

    // query chunk state; lane L owns the chunk's row at position L
    uint32_t qgid = ~0u;      // lane L: the L-th smallest member index of the chunk
    v4i aop[4][2];            // row block rb, K-step ks: row rb * 16 + c16, bytes [64 ks + 16 q, +16)
    int32_t cA[4][4];         // NA' << 6 | position of row rb * 16 + 4 q + j
    int32_t naOwn = 0;        // NA' << 6 of row L
    uint64_t rowBest = kNoMatch;
    int32_t* sR = &sRow[wave][0];
    auto flush_rows = [&]() {
        if (rowBest != kNoMatch && qgid != ~0u)
            atomicMin(reinterpret_cast<unsigned long long*>(resQ + qgid), (unsigned long long)rowBest);
        rowBest = kNoMatch;
    };
    bool newQ = true;
    for (;;) {
        if (newQ) {
            flush_rows();
            const uint32_t i = qt * 64 + (uint32_t)lane;
            const uint32_t g = i < qCnt ? Q.buckets[(size_t)(fb >> 8) * Q.n + qBeg + i] : ~0u;
            qgid = wave_sort_u32(g, lane);
            int32_t na[4];
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                const uint32_t gr = (uint32_t)__shfl((int)qgid, rb * 16 + c16);
                const bool valid = gr != ~0u;
                int32_t nap = 0;
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    // positions past the chunk read member 0 and are zeroed
                    const v4i x = *reinterpret_cast<const v4i*>(qDesc + (size_t)(valid ? gr : 0u) * 128 + ks * 64 + q * 16);
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) {
                        const int xa = x[kk] ^ (int)0x80808080u;
                        nap = __builtin_amdgcn_sdot4(xa, xa, nap, false);
                        aop[rb][ks][kk] = valid ? (x[kk] ^ 0x7F7F7F7F) : 0;
                    }
                }
                nap += __shfl_xor(nap, 16);
                nap += __shfl_xor(nap, 32);
                na[rb] = valid ? nap : kPastChunk;  // lanes with c16 = r hold row rb * 16 + r
            }
#pragma unroll
            for (int rb = 0; rb < 4; rb++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    cA[rb][j] = (__shfl(na[rb], 4 * q + j) << 6) | (rb * 16 + 4 * q + j);
            naOwn = (q == 0 ? na[0] : q == 1 ? na[1] : q == 2 ? na[2] : na[3]) << 6;
        }
        // train chunk
        const uint32_t i = tt * 64 + (uint32_t)lane;
        const uint32_t g = i < tCnt ? T.buckets[(size_t)(fb >> 8) * T.n + tBeg + i] : ~0u;
        const uint32_t tgid = wave_sort_u32(g, lane);
        v4i bop[4][2];
        int32_t nb[4], rB[4];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            const uint32_t gc = (uint32_t)__shfl((int)tgid, cb * 16 + c16);
            const bool valid = gc != ~0u;
            int32_t npb = 0;
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const v4i x = *reinterpret_cast<const v4i*>(tDesc + (size_t)(valid ? gc : 0u) * 128 + ks * 64 + q * 16);
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const int xw = x[kk] ^ (int)0x80808080u;
                    npb = __builtin_amdgcn_sdot4(xw, xw, npb, false);
                    npb = __builtin_amdgcn_sdot4(xw, 0x02020202, npb, false);
                    bop[cb][ks][kk] = valid ? xw : 0;
                }
            }
            npb += __shfl_xor(npb, 16);
            npb += __shfl_xor(npb, 32);
            nb[cb] = valid ? npb : kPastChunk;  // NB' of column cb * 16 + c16
            rB[cb] = (nb[cb] << 6) | (cb * 16 + c16);
        }
        // the 64 rows in two halves of 32 (rb = 2h, 2h + 1): a half's products
        // are folded into the row keys and the column minima before the next
        // half's MFMAs, so only 32 accumulator registers are live
        // row rb * 16 + 4 q + j: minimum over the lane's 4 columns, to sR
        int32_t cm[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};  // column cb * 16 + c16 over the lane's rows
#pragma unroll
        for (int h = 0; h < 2; h++) {
            v4i acc[2][4];
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int r = 0; r < 2; r++)
#pragma unroll
                    for (int cb = 0; cb < 4; cb++)
                        acc[r][cb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aop[2 * h + r][ks], bop[cb][ks],
                                                                          ks == 0 ? v4i{0, 0, 0, 0} : acc[r][cb], 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int rb = 2 * h + r;
                    const int32_t r0 = shl7_add(acc[r][0][j], rB[0]), r1 = shl7_add(acc[r][1][j], rB[1]);
                    const int32_t r2 = shl7_add(acc[r][2][j], rB[2]), r3 = shl7_add(acc[r][3][j], rB[3]);
                    sR[(rb * 16 + 4 * q + j) * kRowStride + c16] = min(min(r0, r1), min(r2, r3));
                }
            if (doCols) {
#pragma unroll
                for (int cb = 0; cb < 4; cb++)
#pragma unroll
                    for (int r = 0; r < 2; r++)
#pragma unroll
                        for (int j = 0; j < 4; j++) cm[cb] = min(cm[cb], shl7_add(acc[r][cb][j], cA[2 * h + r][j]));
            }
        }
        // rows: lane L takes row L's 16 column-class minima back from sR
        // (the wave's own LDS traffic completes in issue order)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        {
            int32_t m = sR[lane * kRowStride];
#pragma unroll
            for (int c = 1; c < 16; c++) m = min(m, sR[lane * kRowStride + c]);
            const int32_t key = m + naOwn;  // distance << 6 | column position of row L
            const uint32_t tg = (uint32_t)__shfl((int)tgid, key & 63);
            if (key < kFoundLimit) {
                const uint64_t cand = make_key((uint32_t)(key >> 6), tg);
                rowBest = cand < rowBest ? cand : rowBest;
            }
        }
        if (doCols) {
            // lane L ends with column L's minimum over the chunk's rows
            const auto sa = __builtin_amdgcn_permlane32_swap((uint32_t)cm[0], (uint32_t)cm[2], false, false);
            const auto sb = __builtin_amdgcn_permlane32_swap((uint32_t)cm[1], (uint32_t)cm[3], false, false);
            const int32_t t02 = min((int32_t)sa[0], (int32_t)sa[1]);
            const int32_t t13 = min((int32_t)sb[0], (int32_t)sb[1]);
            const auto sc = __builtin_amdgcn_permlane16_swap((uint32_t)t02, (uint32_t)t13, false, false);
            const int32_t nbOwn = q == 0 ? nb[0] : q == 1 ? nb[1] : q == 2 ? nb[2] : nb[3];
            const int32_t key = min((int32_t)sc[0], (int32_t)sc[1]) + (nbOwn << 6);
            const uint32_t qg = (uint32_t)__shfl((int)qgid, key & 63);
            if (tgid != ~0u && key < kFoundLimit)
                atomicMin(reinterpret_cast<unsigned long long*>(resT + tgid),
                          (unsigned long long)make_key((uint32_t)(key >> 6), qg));
        }
        // next block
        if (++k >= kEnd) break;
        newQ = ++tt == nT;
        if (newQ) {
            tt = 0;
            if (++qt == nQ) {
                qt = 0;
                do {
                    fb++;
                } while (uniform(sPre[fb + 1]) == uniform(sPre[fb]));
                bucket();
                nT = (tCnt + 63) >> 6;
                nQ = (qCnt + 63) >> 6;
            }
        }
    }
    flush_rows();
}

__global__ void match_init_kernel(const MatchBatch batch)
{
    const uint32_t task = blockIdx.z;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < batch.q[task].n) batch.resQ[task][x] = kNoMatch;
    if (batch.resT[task] != nullptr && x < batch.t.n) batch.resT[task][x] = kNoMatch;
}

// ---------------------------------------------------------------------------
// Config-1 mode (no reference counterpart: the reference computes the 128-bit
// remap hash, computeHash.cu:474-496, but never matches with it; its
// matchWithoutRemap.cu is empty): for every query the candidates are the train
// descriptors sharing its bucket in any of the first nSchemes tables, ranked by
// the Hamming distance of their 128-bit remap hashes (ties: lower index); the
// kKeep first are re-ranked by exact L2^2 (ties: lower index).
//
// One wavefront per query. Lanes stride over the query's bucket in each table,
// reading the bucket-ordered records of the train image (its members' indices
// and remap hashes, gathered once per call). Each lane keeps its kKeep smallest
// distinct keys (hamming << 24 | index) sorted in registers: an item met again
// in a later table has the same key, so it is either still in the list or
// larger than the list's last key, and is not inserted twice. kKeep rounds of a
// wave minimum merge the 64 lists; lanes holding the minimum all drop it, so an
// item met by two lanes is selected once. The exact distances of the kept
// candidates use eight lanes per candidate (16 bytes each, v_dot4 on u8) and a
// wave minimum of (distance << 32 | index).
struct RerankImage {
    const uint8_t* desc;
    const uint64_t* codes;  // 8 code bytes per descriptor
    const uint4* remap;     // 128-bit remap hash per descriptor
    const uint32_t* ends;   // [8][256] inclusive prefix
    const uint32_t* buckets;
    uint32_t n;
    // the members' remap hashes and indices in bucket order, [nSchemes][n]
    // (rerank_recs_kernel), so that lanes striding over a bucket read
    // consecutive entries
    uint4* recRemap;
    uint32_t* recIdx;
};
struct RerankTask {
    RerankImage q, t;  // queries matched against the train image t
    uint64_t* res;     // [q.n] (distance << 32 | train index)
};
constexpr int kRerankMaxTasks = 30;
struct RerankBatch {
    RerankTask t[kRerankMaxTasks];
};
static_assert(sizeof(RerankBatch) + sizeof(uint32_t) <= 4096, "the batch is kept within 4 KiB of kernel arguments");
constexpr int kRerankWaves = 4;

__device__ inline uint32_t wave_min_u32(uint32_t x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o));
    return x;
}
__device__ inline uint64_t wave_min_u64(uint64_t x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = __shfl_xor(x, o);
        x = y < x ? y : x;
    }
    return x;
}

// records of image blockIdx.y (0: the train image, 1 + i: query image i) in
// bucket order, one per (table, position)
struct RecBatch {
    RerankImage im[kRerankMaxTasks + 1];
};
__global__ __launch_bounds__(256) void rerank_recs_kernel(const RecBatch batch, uint32_t nSchemes)
{
    const RerankImage& im = batch.im[blockIdx.y];
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (im.recIdx == nullptr || x >= nSchemes * im.n) return;
    const uint32_t j = im.buckets[x];
    im.recIdx[x] = j;
    im.recRemap[x] = im.remap[j];
}

template <int kKeep>
__global__ __launch_bounds__(kRerankWaves * 64) void rerank_kernel(const RerankBatch batch, uint32_t nSchemes)
{
    const RerankTask& T = batch.t[blockIdx.y];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t qi = blockIdx.x * kRerankWaves + (threadIdx.x >> 6);
    if (qi >= T.q.n) return;
    const uint64_t qc = T.q.codes[qi];
    const uint4 qr = T.q.remap[qi];
    uint32_t top[kKeep];
#pragma unroll
    for (int k = 0; k < kKeep; k++) top[k] = ~0u;
    for (uint32_t s = 0; s < nSchemes; s++) {
        const uint32_t code = (uint32_t)(qc >> (8 * s)) & 0xFF;
        const uint32_t end = T.t.ends[s * 256 + code];
        const uint32_t beg = code == 0 ? 0u : T.t.ends[s * 256 + code - 1];
        const uint4* recR = T.t.recRemap + (size_t)s * T.t.n;
        const uint32_t* recI = T.t.recIdx + (size_t)s * T.t.n;
        // four records per lane in flight: their loads are issued before the
        // first is tested
        for (uint32_t p0 = beg + lane; p0 < end; p0 += 4 * 64) {
            uint32_t jj[4];
            uint4 rr[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t p = p0 + u * 64;
                jj[u] = p < end ? recI[p] : ~0u;
                rr[u] = p < end ? recR[p] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (p0 + u * 64 >= end) continue;
                const uint4 r = rr[u];
                const uint32_t h = (uint32_t)(__popc(r.x ^ qr.x) + __popc(r.y ^ qr.y) + __popc(r.z ^ qr.z) +
                                              __popc(r.w ^ qr.w));
                uint32_t key = (h << 24) | jj[u];
                if (key >= top[kKeep - 1]) continue;
                // an item met in an earlier table (same key) is listed once
                bool dup = false;
#pragma unroll
                for (int k = 0; k < kKeep; k++) dup = dup || top[k] == key;
                if (dup) continue;
#pragma unroll
                for (int k = 0; k < kKeep; k++) {
                    const uint32_t lo = min(top[k], key);
                    key = max(top[k], key);
                    top[k] = lo;
                }
            }
        }
    }
    // the kKeep smallest keys of the wave, in order
    uint32_t sel[kKeep];
#pragma unroll
    for (int k = 0; k < kKeep; k++) {
        const uint32_t m = wave_min_u32(top[0]);
        sel[k] = m;
        if (top[0] == m && m != ~0u) {
#pragma unroll
            for (int u = 0; u + 1 < kKeep; u++) top[u] = top[u + 1];
            top[kKeep - 1] = ~0u;
        }
    }
    // exact L2^2 of the kept candidates: lane = 8 x candidate slot + 16-byte part
    const uint32_t part = lane & 7, slot = lane >> 3;
    const uint4 a = *reinterpret_cast<const uint4*>(T.q.desc + (size_t)qi * 128 + part * 16);
    uint32_t aa = 0;
    aa = __builtin_amdgcn_udot4(a.x, a.x, aa, false);
    aa = __builtin_amdgcn_udot4(a.y, a.y, aa, false);
    aa = __builtin_amdgcn_udot4(a.z, a.z, aa, false);
    aa = __builtin_amdgcn_udot4(a.w, a.w, aa, false);
    uint64_t best = kNoMatch;
#pragma unroll
    for (int g = 0; g < kKeep / 8; g++) {
        uint32_t key = sel[0];
#pragma unroll
        for (int k = 1; k < 8; k++)
            if ((uint32_t)k == slot) key = sel[g * 8 + k];
        if (slot == 0) key = sel[g * 8];
        const bool valid = key != ~0u;
        const uint32_t j = key & 0xFFFFFFu;
        int32_t d = 0;
        if (valid) {
            const uint4 b = *reinterpret_cast<const uint4*>(T.t.desc + (size_t)j * 128 + part * 16);
            uint32_t bb = 0, ab = 0;
            bb = __builtin_amdgcn_udot4(b.x, b.x, bb, false);
            bb = __builtin_amdgcn_udot4(b.y, b.y, bb, false);
            bb = __builtin_amdgcn_udot4(b.z, b.z, bb, false);
            bb = __builtin_amdgcn_udot4(b.w, b.w, bb, false);
            ab = __builtin_amdgcn_udot4(a.x, b.x, ab, false);
            ab = __builtin_amdgcn_udot4(a.y, b.y, ab, false);
            ab = __builtin_amdgcn_udot4(a.z, b.z, ab, false);
            ab = __builtin_amdgcn_udot4(a.w, b.w, ab, false);
            d = (int32_t)(aa + bb) - 2 * (int32_t)ab;
        }
        d += __shfl_xor(d, 1);
        d += __shfl_xor(d, 2);
        d += __shfl_xor(d, 4);
        const uint64_t k64 = valid ? make_key((uint32_t)d, j) : kNoMatch;
        best = k64 < best ? k64 : best;
    }
    best = wave_min_u64(best);
    if (lane == 0) T.res[qi] = best;
}

}  // namespace cashash
}  // namespace rsfm_amd

using namespace rsfm_amd;

extern "C" int rsfm_cashash_accumulate_desc(uint64_t* acc, uint64_t* sqrAcc, int initAcc, const void* desc, uint32_t nbDesc,
                                            hipStream_t stream)
{
    RSFM_RETURN_IF(acc == nullptr, RSFM_ERR_INVALID, "accumulate_desc: acc is null");
    RSFM_RETURN_IF(nbDesc > 0 && (desc == nullptr || reinterpret_cast<uintptr_t>(desc) % 16), RSFM_ERR_INVALID,
                   "accumulate_desc: desc must be non-null and 16-byte aligned");
    if (initAcc) {
        cashash::zero_u64_kernel<<<1, 128, 0, stream>>>(acc, 128);
        if (sqrAcc) cashash::zero_u64_kernel<<<1, 128, 0, stream>>>(sqrAcc, 128);
        RSFM_CHECK(check_launch("accumulate_desc init"));
    }
    if (nbDesc == 0) return RSFM_OK;
    cashash::accumulate_kernel<<<div_up(nbDesc, cashash::kAccDescPerBlock), cashash::kAccThreads, 0, stream>>>(
        acc, sqrAcc, static_cast<const uint8_t*>(desc), nbDesc);
    return check_launch("accumulate_desc");
}

extern "C" int rsfm_cashash_compute_bias(int32_t* bias, const int8_t* weights, const int32_t* biasBase, const uint64_t* acc,
                                         uint64_t nbDesc, uint32_t nbChannels, hipStream_t stream)
{
    RSFM_RETURN_IF(bias == nullptr || weights == nullptr || biasBase == nullptr || acc == nullptr, RSFM_ERR_INVALID,
                   "compute_bias: null pointer");
    RSFM_RETURN_IF(nbChannels != 64 && nbChannels != 192, RSFM_ERR_INVALID, "compute_bias: nbChannels must be 64 or 192");
    RSFM_RETURN_IF(nbDesc == 0, RSFM_ERR_INVALID, "compute_bias: nbDesc must be > 0");
    cashash::bias_kernel<<<nbChannels / 64, 512, 0, stream>>>(bias, weights, biasBase, acc, nbDesc, nbChannels);
    return check_launch("compute_bias");
}

static int launch_hash(const int8_t* weights, const int32_t* bias, const void* desc, uint32_t nbDesc, uint8_t* bucketIndices,
                       uint32_t* remap, uint32_t* histogram, hipStream_t stream, bool withRemap)
{
    RSFM_RETURN_IF(weights == nullptr || bias == nullptr, RSFM_ERR_INVALID, "compute_hash: weights/bias null");
    RSFM_RETURN_IF(nbDesc > 0 && (desc == nullptr || bucketIndices == nullptr), RSFM_ERR_INVALID,
                   "compute_hash: desc/bucketIndices null");
    RSFM_RETURN_IF(withRemap && nbDesc > 0 && remap == nullptr, RSFM_ERR_INVALID, "compute_hash: remap null");
    RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(desc) | reinterpret_cast<uintptr_t>(weights)) % 16 ||
                       reinterpret_cast<uintptr_t>(bucketIndices) % 8 || reinterpret_cast<uintptr_t>(remap) % 16,
                   RSFM_ERR_INVALID, "compute_hash: misaligned pointer");
    if (histogram) {
        cashash::zero_u32_kernel<<<8, 256, 0, stream>>>(histogram, 2048);
        RSFM_CHECK(check_launch("compute_hash histogram init"));
    }
    if (nbDesc == 0) return RSFM_OK;
    const uint32_t blocks = div_up(nbDesc, 32 * cashash::kHashWaves);
    uint32_t* codes = reinterpret_cast<uint32_t*>(bucketIndices);
    if (withRemap)
        cashash::hash_kernel<6><<<blocks, cashash::kHashWaves * 64, 0, stream>>>(
            weights, bias, static_cast<const uint8_t*>(desc), nbDesc, codes, remap, histogram);
    else
        cashash::hash_kernel<2><<<blocks, cashash::kHashWaves * 64, 0, stream>>>(
            weights, bias, static_cast<const uint8_t*>(desc), nbDesc, codes, nullptr, histogram);
    return check_launch("compute_hash");
}

extern "C" int rsfm_cashash_compute_hash_bucket_only(const int8_t* weights, const int32_t* bias, const void* desc,
                                                     uint32_t nbDesc, uint8_t* bucketIndices, uint32_t* histogram,
                                                     hipStream_t stream)
{
    return launch_hash(weights, bias, desc, nbDesc, bucketIndices, nullptr, histogram, stream, false);
}

extern "C" int rsfm_cashash_compute_hash_bucket_and_remap(const int8_t* weights, const int32_t* bias, const void* desc,
                                                          uint32_t nbDesc, uint8_t* bucketIndices, uint32_t* remap,
                                                          uint32_t* histogram, hipStream_t stream)
{
    return launch_hash(weights, bias, desc, nbDesc, bucketIndices, remap, histogram, stream, true);
}

extern "C" int rsfm_cashash_build_histogram(uint32_t* histogram, const uint8_t* bucketIndices, uint32_t nbDesc,
                                            hipStream_t stream)
{
    RSFM_RETURN_IF(histogram == nullptr, RSFM_ERR_INVALID, "build_histogram: histogram null");
    RSFM_RETURN_IF(nbDesc > 0 && (bucketIndices == nullptr || reinterpret_cast<uintptr_t>(bucketIndices) % 8),
                   RSFM_ERR_INVALID, "build_histogram: bucketIndices null or misaligned");
    cashash::zero_u32_kernel<<<8, 256, 0, stream>>>(histogram, 2048);
    RSFM_CHECK(check_launch("build_histogram init"));
    if (nbDesc == 0) return RSFM_OK;
    cashash::histogram_kernel<<<div_up(nbDesc, cashash::kHistPerBlock), cashash::kHistThreads, 0, stream>>>(
        histogram, bucketIndices, nbDesc);
    return check_launch("build_histogram");
}

extern "C" int rsfm_cashash_compute_bucket_bound(uint32_t* lower, uint32_t* upper, uint32_t* maxBucketSize,
                                                 const uint32_t* histogram, hipStream_t stream)
{
    RSFM_RETURN_IF(lower == nullptr || upper == nullptr || histogram == nullptr, RSFM_ERR_INVALID,
                   "compute_bucket_bound: null pointer");
    cashash::bounds_kernel<<<1, cashash::kBoundsThreads, 0, stream>>>(lower, upper, maxBucketSize, histogram);
    return check_launch("compute_bucket_bound");
}

extern "C" int rsfm_cashash_build_buckets(const uint8_t* bucketIndices, uint32_t nbDesc, uint32_t* idxAllocator,
                                          uint32_t* buckets, hipStream_t stream)
{
    if (nbDesc == 0) return RSFM_OK;
    RSFM_RETURN_IF(bucketIndices == nullptr || idxAllocator == nullptr || buckets == nullptr, RSFM_ERR_INVALID,
                   "build_buckets: null pointer");
    cashash::buckets_kernel<<<div_up(nbDesc, cashash::kBuildPerBlock), cashash::kBuildThreads, 0, stream>>>(
        bucketIndices, nbDesc, idxAllocator, buckets);
    return check_launch("build_buckets");
}

extern "C" int rsfm_cashash_find_best_match(const rsfm_cashash_image* queries, const uint32_t* querySize,
                                            const rsfm_cashash_image* train, uint32_t trainSize,
                                            const rsfm_cashash_result* results, uint32_t nbTasks,
                                            const uint32_t* maxQueryBucketSize, uint32_t maxTrainBucketSize,
                                            uint32_t nbSchemes, hipStream_t stream)
{
    (void)maxQueryBucketSize;
    (void)maxTrainBucketSize;
    using namespace cashash;
    RSFM_RETURN_IF(nbTasks > 0 && (queries == nullptr || querySize == nullptr || results == nullptr || train == nullptr),
                   RSFM_ERR_INVALID, "find_best_match: null pointer");
    RSFM_RETURN_IF(nbSchemes == 0 || nbSchemes > kSchemes, RSFM_ERR_INVALID, "find_best_match: nbSchemes must be 1..8");
    for (uint32_t b = 0; b < nbTasks; b += kMatchMaxTasks) {
        const uint32_t n = std::min<uint32_t>(kMatchMaxTasks, nbTasks - b);
        MatchBatch batch{};
        batch.t = DevImage{static_cast<const uint8_t*>(train->desc), train->bucketEnds, train->buckets, trainSize};
        uint32_t maxN = trainSize;
        for (uint32_t i = 0; i < n; i++) {
            const rsfm_cashash_image& q = queries[b + i];
            RSFM_RETURN_IF(querySize[b + i] > 0 && (q.desc == nullptr || q.bucketEnds == nullptr || q.buckets == nullptr ||
                                                    results[b + i].query == nullptr),
                           RSFM_ERR_INVALID, "find_best_match: query image incomplete");
            batch.q[i] = DevImage{static_cast<const uint8_t*>(q.desc), q.bucketEnds, q.buckets, querySize[b + i]};
            batch.resQ[i] = reinterpret_cast<uint64_t*>(results[b + i].query);
            batch.resT[i] = reinterpret_cast<uint64_t*>(results[b + i].train);
            maxN = std::max(maxN, querySize[b + i]);
        }
        if (maxN == 0) continue;
        match_init_kernel<<<dim3(div_up(maxN, 256), 1, n), 256, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("find_best_match init"));
        // about 512 workgroups (2048 waves) per launch, each wave a few
        // 64 x 64 blocks at 8K keypoints
        const uint32_t perTask = std::max<uint32_t>(8u, div_up(512u, n));
        match_kernel<<<dim3(perTask, 1, n), kMatchWaves * 64, 0, stream>>>(batch, nbSchemes);
        RSFM_CHECK(check_launch("find_best_match"));
    }
    return RSFM_OK;
}

extern "C" int rsfm_cashash_find_best_match_rerank(const rsfm_cashash_rerank_image* queries,
                                                   const rsfm_cashash_rerank_image* train,
                                                   const rsfm_cashash_result* results, uint32_t nbTasks,
                                                   uint32_t nbSchemes, uint32_t keep, hipStream_t stream)
{
    using namespace cashash;
    RSFM_RETURN_IF(nbTasks > 0 && (queries == nullptr || train == nullptr || results == nullptr), RSFM_ERR_INVALID,
                   "find_best_match_rerank: null pointer");
    RSFM_RETURN_IF(nbSchemes == 0 || nbSchemes > kSchemes, RSFM_ERR_INVALID,
                   "find_best_match_rerank: nbSchemes must be 1..8");
    RSFM_RETURN_IF(keep != 8 && keep != 16 && keep != 32, RSFM_ERR_INVALID,
                   "find_best_match_rerank: keep must be 8, 16 or 32");
    auto image = [](const rsfm_cashash_rerank_image& im, RerankImage& out) -> int {
        RSFM_RETURN_IF(im.nbDesc > 0 && (im.desc == nullptr || im.codes == nullptr || im.remap == nullptr ||
                                         im.bucketEnds == nullptr || im.buckets == nullptr),
                       RSFM_ERR_INVALID, "find_best_match_rerank: image incomplete");
        RSFM_RETURN_IF(im.nbDesc > (1u << 24), RSFM_ERR_INVALID,
                       "find_best_match_rerank: at most 2^24 descriptors per image");
        RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(im.desc) | reinterpret_cast<uintptr_t>(im.remap)) % 16 ||
                           reinterpret_cast<uintptr_t>(im.codes) % 8,
                       RSFM_ERR_INVALID, "find_best_match_rerank: misaligned pointer");
        out = RerankImage{static_cast<const uint8_t*>(im.desc), reinterpret_cast<const uint64_t*>(im.codes),
                          reinterpret_cast<const uint4*>(im.remap), im.bucketEnds, im.buckets, im.nbDesc};
        return RSFM_OK;
    };
    std::vector<RerankImage> imgs(1 + (size_t)nbTasks);  // the train image, then the query images
    RSFM_CHECK(image(*train, imgs[0]));
    for (uint32_t i = 0; i < nbTasks; i++) {
        RSFM_CHECK(image(queries[i], imgs[1 + i]));
        RSFM_RETURN_IF(imgs[1 + i].n > 0 && results[i].query == nullptr, RSFM_ERR_INVALID,
                       "find_best_match_rerank: results.query is null");
    }
    // every image's members in bucket order (20 B per member and table), from
    // the stream-ordered allocator
    size_t nRecs = 0;
    for (const auto& im : imgs) nRecs += (size_t)nbSchemes * im.n;
    if (nRecs == 0) return RSFM_OK;
    void* ws = nullptr;
    RSFM_CHECK(check_hip(hipMallocAsync(&ws, nRecs * (sizeof(uint4) + sizeof(uint32_t)), stream),
                         "find_best_match_rerank: records"));
    uint4* rr = static_cast<uint4*>(ws);
    uint32_t* ri = reinterpret_cast<uint32_t*>(rr + nRecs);
    for (auto& im : imgs) {
        im.recRemap = im.n > 0 ? rr : nullptr;
        im.recIdx = im.n > 0 ? ri : nullptr;
        rr += (size_t)nbSchemes * im.n;
        ri += (size_t)nbSchemes * im.n;
    }
    int rc = RSFM_OK;
    for (size_t b = 0; b < imgs.size() && rc == RSFM_OK; b += kRerankMaxTasks + 1) {
        const size_t n = std::min<size_t>(kRerankMaxTasks + 1, imgs.size() - b);
        RecBatch rb{};
        uint32_t maxN = 0;
        for (size_t i = 0; i < n; i++) {
            rb.im[i] = imgs[b + i];
            maxN = std::max(maxN, imgs[b + i].n);
        }
        if (maxN == 0) continue;
        rerank_recs_kernel<<<dim3(div_up(nbSchemes * maxN, 256), (uint32_t)n), 256, 0, stream>>>(rb, nbSchemes);
        rc = check_launch("find_best_match_rerank records");
    }
    // every task's query direction, then (where asked) its train direction
    std::vector<RerankTask> all;
    const RerankImage& t = imgs[0];
    for (uint32_t i = 0; i < nbTasks; i++) {
        const RerankImage& q = imgs[1 + i];
        if (q.n > 0) all.push_back(RerankTask{q, t, reinterpret_cast<uint64_t*>(results[i].query)});
        if (results[i].train != nullptr && t.n > 0)
            all.push_back(RerankTask{t, q, reinterpret_cast<uint64_t*>(results[i].train)});
    }
    for (size_t b = 0; b < all.size() && rc == RSFM_OK; b += kRerankMaxTasks) {
        const size_t n = std::min<size_t>(kRerankMaxTasks, all.size() - b);
        RerankBatch batch{};
        uint32_t maxN = 0;
        for (size_t i = 0; i < n; i++) {
            batch.t[i] = all[b + i];
            maxN = std::max(maxN, batch.t[i].q.n);
        }
        const dim3 grid(div_up(maxN, kRerankWaves), (uint32_t)n);
        if (keep == 8) rerank_kernel<8><<<grid, kRerankWaves * 64, 0, stream>>>(batch, nbSchemes);
        else if (keep == 16) rerank_kernel<16><<<grid, kRerankWaves * 64, 0, stream>>>(batch, nbSchemes);
        else rerank_kernel<32><<<grid, kRerankWaves * 64, 0, stream>>>(batch, nbSchemes);
        rc = check_launch("find_best_match_rerank");
    }
    const int rf = check_hip(hipFreeAsync(ws, stream), "find_best_match_rerank: records release");
    return rc != RSFM_OK ? rc : rf;
}
