// bf_aux.hip - descriptor precompute, the vector-ALU matcher and match filtering.
#include "internal.hpp"
#include <algorithm>
#include <vector>

namespace rsfm_amd {
namespace aux {

// sum of squares, 8 lanes x 16 B per descriptor (kernelPreCompSift8bSqrNorm,
// bruteForceMatchIMMA.cu:704-731)
__global__ void sqrnorm_u8_kernel(uint32_t* __restrict__ out, const uint8_t* __restrict__ desc, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = t >> 3, part = t & 7;
    uint32_t acc = 0;
    if (d < n) {
        const uint4 x = *reinterpret_cast<const uint4*>(desc + (size_t)d * 128 + part * 16);
        acc = __builtin_amdgcn_udot4(x.x, x.x, acc, false);
        acc = __builtin_amdgcn_udot4(x.y, x.y, acc, false);
        acc = __builtin_amdgcn_udot4(x.z, x.z, acc, false);
        acc = __builtin_amdgcn_udot4(x.w, x.w, acc, false);
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    acc += __shfl_xor(acc, 4);
    if (d < n && part == 0) out[d] = acc;
}

// 4-bit descriptor (64 B): 4 lanes x 16 B (kernelPreCompSift4bSqrNorm, :671-701)
__global__ void sqrnorm_u4_kernel(uint32_t* __restrict__ out, const uint8_t* __restrict__ desc, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = t >> 2, part = t & 3;
    uint32_t acc = 0;
    if (d < n) {
        const uint4 x = *reinterpret_cast<const uint4*>(desc + (size_t)d * 64 + part * 16);
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t lo = w[q] & 0x0F0F0F0Fu, hn = (w[q] >> 4) & 0x0F0F0F0Fu;
            acc = __builtin_amdgcn_udot4(lo, lo, acc, false);
            acc = __builtin_amdgcn_udot4(hn, hn, acc, false);
        }
    }
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (d < n && part == 0) out[d] = acc;
}

// high-nibble pack, one lane per 16 source bytes (kernelSiftDesc8bTo4b, :637-660)
__global__ void desc8b_to_4b_kernel(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = t >> 3, part = t & 7;
    if (d >= n) return;
    const uint4 x = *reinterpret_cast<const uint4*>(src + (size_t)d * 128 + part * 16);
    uint2 y;
    y.x = ((x.x & 0xF0F0F0F0u) >> 4) | (x.y & 0xF0F0F0F0u);
    y.y = ((x.z & 0xF0F0F0F0u) >> 4) | (x.w & 0xF0F0F0F0u);
    *reinterpret_cast<uint2*>(dst + (size_t)d * 64 + part * 8) = y;
}

// ---------------------------------------------------------------------------
// Vector-ALU matcher (the IDP path): one query row per lane held in
// registers, train descriptors staged through LDS 64 at a time, u8 dot
// products with v_dot4_u32_u8, L2^2 = |a|^2 + |b|^2 - 2 a.b.
constexpr int kIdpThreads = 256;
constexpr int kIdpTile = 64;
constexpr int kIdpMaxTasks = 32;

struct IdpTask {
    uint32_t nA, nB;
    const uint8_t* A;
    const uint8_t* B;
    uint64_t* bestA;
    uint64_t* bestB;
};
struct IdpBatch {
    IdpTask t[kIdpMaxTasks];
};

__global__ __launch_bounds__(kIdpThreads) void idp_kernel(const IdpBatch batch)
{
    const IdpTask& T = batch.t[blockIdx.z];
    const uint32_t rowBase = blockIdx.x * kIdpThreads;
    if (rowBase >= T.nA || T.nB == 0) return;
    __shared__ uint32_t sB[kIdpTile][33];  // padded rows
    __shared__ uint32_t sNB[kIdpTile];
    __shared__ unsigned long long sColBest[kIdpTile];
    const uint32_t row = rowBase + threadIdx.x;
    const bool valid = row < T.nA;
    uint32_t a[32];
    uint32_t na = 0;
#pragma unroll
    for (int q = 0; q < 32; q++) {
        a[q] = valid ? reinterpret_cast<const uint32_t*>(T.A + (size_t)row * 128)[q] : 0u;
        na = __builtin_amdgcn_udot4(a[q], a[q], na, false);
    }
    uint64_t best = kNoMatch;
    for (uint32_t c0 = 0; c0 < T.nB; c0 += kIdpTile) {
        __syncthreads();
        // stage 64 descriptors: 256 lanes x 32 B
        {
            const uint32_t col = threadIdx.x >> 2, part = threadIdx.x & 3;
            const uint32_t g = c0 + col;
#pragma unroll
            for (int q = 0; q < 8; q++)
                sB[col][part * 8 + q] = g < T.nB ? reinterpret_cast<const uint32_t*>(T.B + (size_t)g * 128)[part * 8 + q] : 0u;
            if (threadIdx.x < kIdpTile) sColBest[threadIdx.x] = kNoMatch;
        }
        __syncthreads();
        if (threadIdx.x < kIdpTile) {
            uint32_t nb = 0;
#pragma unroll
            for (int q = 0; q < 32; q++) nb = __builtin_amdgcn_udot4(sB[threadIdx.x][q], sB[threadIdx.x][q], nb, false);
            sNB[threadIdx.x] = nb;
        }
        __syncthreads();
        const uint32_t ncol = min((uint32_t)kIdpTile, T.nB - c0);
        for (uint32_t j = 0; j < ncol; j++) {
            uint32_t dot = 0;
#pragma unroll
            for (int q = 0; q < 32; q++) dot = __builtin_amdgcn_udot4(a[q], sB[j][q], dot, false);
            const uint32_t dist = na + sNB[j] - 2u * dot;
            if (valid) {
                best = min(best, make_key(dist, c0 + j));
                if (T.bestB) atomicMin(&sColBest[j], (unsigned long long)make_key(dist, row));
            }
        }
        __syncthreads();
        if (T.bestB && threadIdx.x < ncol && sColBest[threadIdx.x] != kNoMatch)
            atomicMin(reinterpret_cast<unsigned long long*>(T.bestB + c0 + threadIdx.x), sColBest[threadIdx.x]);
    }
    if (valid) T.bestA[row] = best;
}

__global__ void idp_init_kernel(const IdpBatch batch)
{
    const IdpTask& T = batch.t[blockIdx.z];
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < T.nA) T.bestA[x] = kNoMatch;
    if (T.bestB != nullptr && x < T.nB) T.bestB[x] = kNoMatch;
}

// ---------------------------------------------------------------------------
// Cross-check + ratio-test filter with in-order compaction, one workgroup.
constexpr int kFilterThreads = 1024;

constexpr int kFilterMaxTasks = 32;
typedef ByValue<FilterTask, kFilterMaxTasks> FilterBatch;

// Queries are taken in rounds of kFilterThreads consecutive indices (thread t
// takes index round * kFilterThreads + t, so every load is coalesced) and
// kFilterRounds rounds per trip: all their loads and gathers are issued
// before the first flag is needed. A kept query's slot is its rank in its
// wave's ballot (v_mbcnt) plus the kept counts of all (round, wave) pairs
// before its own, so the output stays in ascending query order; three
// barriers per trip.
constexpr int kFilterRounds = 8;
constexpr int kFilterWaves = kFilterThreads / 64;

__device__ void filter_body(const FilterTask& T, uint32_t num, uint32_t den)
{
    __shared__ uint32_t sCnt[kFilterRounds * kFilterWaves];
    __shared__ uint32_t sPre[kFilterRounds * kFilterWaves];
    __shared__ uint32_t sTot;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nQ = T.nQ;
    const uint64_t nn = (uint64_t)num * num, dd = (uint64_t)den * den;
    uint32_t base = 0;  // kept matches of the earlier trips (the same in every thread)
    for (uint32_t r0 = 0; r0 < nQ; r0 += kFilterRounds * kFilterThreads) {
        uint64_t f[kFilterRounds];
        uint32_t back[kFilterRounds];
        uint64_t d2[kFilterRounds];
#pragma unroll
        for (int k = 0; k < kFilterRounds; k++) {
            const uint32_t i = r0 + k * kFilterThreads + tid;
            f[k] = i < nQ ? T.fwd[i] : kNoMatch;
            d2[k] = (T.second != nullptr && i < nQ) ? T.second[i] : 0;
        }
#pragma unroll
        for (int k = 0; k < kFilterRounds; k++) {
            const uint32_t i = r0 + k * kFilterThreads + tid;
            const uint32_t t = (uint32_t)f[k];
            back[k] = (T.bwd != nullptr && t < T.nT) ? (uint32_t)T.bwd[t] : i;
        }
        uint32_t keepMask = 0;
#pragma unroll
        for (int k = 0; k < kFilterRounds; k++) {
            const uint32_t i = r0 + k * kFilterThreads + tid;
            const uint32_t t = (uint32_t)f[k];
            bool keep = i < nQ && t != ~0u && t < T.nT && back[k] == i;
            if (T.second != nullptr) keep = keep && dd * (f[k] >> 32) < nn * (d2[k] >> 32);
            keepMask |= (uint32_t)keep << k;
        }
        float4 pt[kFilterRounds];
        if (T.pts) {
#pragma unroll
            for (int k = 0; k < kFilterRounds; k++) {
                if ((keepMask >> k) & 1) {
                    const uint32_t i = r0 + k * kFilterThreads + tid;
                    const float2 a = T.kpQ[i], b = T.kpT[(uint32_t)f[k]];
                    pt[k] = make_float4(a.x, a.y, b.x, b.y);
                }
            }
        }
        // rank: kept counts per (round, wave) into LDS, one wave turns them
        // into exclusive offsets in (round, wave) order, i.e. query order
        uint32_t below[kFilterRounds];
#pragma unroll
        for (int k = 0; k < kFilterRounds; k++) {
            const uint64_t m = __ballot((keepMask >> k) & 1);
            below[k] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (lane == 0) sCnt[k * kFilterWaves + wave] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (wave == 0) {
            constexpr int kPerLane = kFilterRounds * kFilterWaves / 64;
            uint32_t c[kPerLane], sum = 0;
#pragma unroll
            for (int j = 0; j < kPerLane; j++) {
                c[j] = sCnt[lane * kPerLane + j];
                sum += c[j];
            }
            uint32_t incl = sum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(incl, o);
                if (lane >= (uint32_t)o) incl += v;
            }
            uint32_t run = incl - sum;
#pragma unroll
            for (int j = 0; j < kPerLane; j++) {
                sPre[lane * kPerLane + j] = run;
                run += c[j];
            }
            if (lane == 63) sTot = incl;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kFilterRounds; k++) {
            if (!((keepMask >> k) & 1)) continue;
            const uint32_t pos = base + sPre[k * kFilterWaves + wave] + below[k];
            const uint32_t i = r0 + k * kFilterThreads + tid;
            if (T.out) {
                rsfm_valid_match mv;
                mv.idxQuery = i;
                mv.idxTrain = (uint32_t)f[k];
                mv.distance = (float)(uint32_t)(f[k] >> 32);
                T.out[pos] = mv;
            }
            if (T.pts) T.pts[pos] = pt[k];
        }
        base += sTot;
        __syncthreads();  // sCnt, sPre and sTot are rewritten by the next trip
    }
    if (tid == 0) *T.count = base;
}

// one workgroup per task
template <class Src>
__global__ __launch_bounds__(kFilterThreads) void filter_batch_kernel(const Src batch, uint32_t num, uint32_t den)
{
    filter_body(batch[blockIdx.x], num, den);
}

__global__ __launch_bounds__(kFilterThreads) void filter_kernel(const uint64_t* __restrict__ fwd,
                                                                const uint64_t* __restrict__ second, uint32_t nQ,
                                                                const uint64_t* __restrict__ bwd, uint32_t nT,
                                                                uint32_t num, uint32_t den,
                                                                rsfm_valid_match* __restrict__ out,
                                                                uint32_t* __restrict__ count)
{
    FilterTask T{};
    T.fwd = fwd;
    T.second = second;
    T.nQ = nQ;
    T.bwd = bwd;
    T.nT = nT;
    T.out = out;
    T.count = count;
    filter_body(T, num, den);
}

int launch_filter_table(const FilterTask* tasks, uint32_t nTasks, uint32_t num, uint32_t den, hipStream_t stream)
{
    if (nTasks == 0) return RSFM_OK;
    filter_batch_kernel<<<nTasks, kFilterThreads, 0, stream>>>(InTable<FilterTask>{tasks}, num, den);
    return check_launch("filter table");
}

}  // namespace aux
}  // namespace rsfm_amd

using namespace rsfm_amd;

extern "C" int rsfm_sqrnorm_u8(uint32_t* sqrNorm, const void* desc, uint32_t nbDesc, hipStream_t stream)
{
    if (nbDesc == 0) return RSFM_OK;
    RSFM_RETURN_IF(sqrNorm == nullptr || desc == nullptr, RSFM_ERR_INVALID, "sqrnorm: null pointer");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(desc) % 16, RSFM_ERR_INVALID, "sqrnorm: desc must be 16-byte aligned");
    aux::sqrnorm_u8_kernel<<<div_up(nbDesc * 8, 256), 256, 0, stream>>>(sqrNorm, static_cast<const uint8_t*>(desc), nbDesc);
    return check_launch("sqrnorm_u8");
}

extern "C" int rsfm_sqrnorm_u4(uint32_t* sqrNorm, const void* desc4, uint32_t nbDesc, hipStream_t stream)
{
    if (nbDesc == 0) return RSFM_OK;
    RSFM_RETURN_IF(sqrNorm == nullptr || desc4 == nullptr, RSFM_ERR_INVALID, "sqrnorm_u4: null pointer");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(desc4) % 16, RSFM_ERR_INVALID, "sqrnorm_u4: desc must be 16-byte aligned");
    aux::sqrnorm_u4_kernel<<<div_up(nbDesc * 4, 256), 256, 0, stream>>>(sqrNorm, static_cast<const uint8_t*>(desc4), nbDesc);
    return check_launch("sqrnorm_u4");
}

extern "C" int rsfm_desc8b_to_4b(void* dst, const void* src, uint32_t nbDesc, hipStream_t stream)
{
    if (nbDesc == 0) return RSFM_OK;
    RSFM_RETURN_IF(dst == nullptr || src == nullptr, RSFM_ERR_INVALID, "desc8b_to_4b: null pointer");
    RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(src) % 16) || (reinterpret_cast<uintptr_t>(dst) % 8), RSFM_ERR_INVALID,
                   "desc8b_to_4b: src must be 16-byte and dst 8-byte aligned");
    aux::desc8b_to_4b_kernel<<<div_up(nbDesc * 8, 256), 256, 0, stream>>>(static_cast<uint8_t*>(dst),
                                                                           static_cast<const uint8_t*>(src), nbDesc);
    return check_launch("desc8b_to_4b");
}

extern "C" int rsfm_bf_match_idp_u8(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream)
{
    using namespace aux;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "bf idp: tasks is null");
    for (size_t b = 0; b < nbTasks; b += kIdpMaxTasks) {
        const size_t n = std::min<size_t>(kIdpMaxTasks, nbTasks - b);
        IdpBatch batch{};
        uint32_t maxA = 0, maxB = 0;
        for (size_t i = 0; i < n; i++) {
            const rsfm_bf_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.nbDescA > 0 && (t.descA == nullptr || t.bestMatchA == nullptr), RSFM_ERR_INVALID,
                           "bf idp: descA/bestMatchA must be set when nbDescA > 0");
            RSFM_RETURN_IF(t.nbDescB > 0 && t.descB == nullptr, RSFM_ERR_INVALID, "bf idp: descB must be set");
            RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(t.descA) | reinterpret_cast<uintptr_t>(t.descB)) % 4,
                           RSFM_ERR_INVALID, "bf idp: descriptors must be 4-byte aligned");
            batch.t[i] = IdpTask{t.nbDescA, t.nbDescB, static_cast<const uint8_t*>(t.descA),
                                 static_cast<const uint8_t*>(t.descB), reinterpret_cast<uint64_t*>(t.bestMatchA),
                                 reinterpret_cast<uint64_t*>(t.bestMatchB)};
            maxA = std::max(maxA, t.nbDescA);
            maxB = std::max(maxB, t.nbDescB);
        }
        const uint32_t initX = div_up(std::max(maxA, maxB), 256);
        if (initX == 0) continue;
        idp_init_kernel<<<dim3(initX, 1, (uint32_t)n), 256, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("bf idp init"));
        if (maxA == 0 || maxB == 0) continue;
        idp_kernel<<<dim3(div_up(maxA, kIdpThreads), 1, (uint32_t)n), kIdpThreads, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("bf idp"));
    }
    return RSFM_OK;
}

extern "C" int rsfm_filter_matches_workspace_size(uint32_t nbQueryDesc, size_t* bytes)
{
    (void)nbQueryDesc;
    RSFM_RETURN_IF(bytes == nullptr, RSFM_ERR_INVALID, "bytes is null");
    *bytes = 0;
    return RSFM_OK;
}

extern "C" int rsfm_filter_matches(const rsfm_best_match* fwd, const rsfm_best_match* second, uint32_t nbQueryDesc,
                                   const rsfm_best_match* bwd, uint32_t nbTrainDesc, uint32_t ratioNum,
                                   uint32_t ratioDen, rsfm_valid_match* out, uint32_t* count, void* workspace,
                                   size_t workspaceBytes, hipStream_t stream)
{
    (void)workspace;
    (void)workspaceBytes;
    RSFM_RETURN_IF(count == nullptr, RSFM_ERR_INVALID, "filter: count is null");
    RSFM_RETURN_IF(nbQueryDesc > 0 && (fwd == nullptr || out == nullptr), RSFM_ERR_INVALID, "filter: null pointer");
    RSFM_RETURN_IF(second != nullptr && (ratioNum == 0 || ratioDen == 0), RSFM_ERR_INVALID,
                   "filter: ratio must be num/den with both > 0");
    aux::filter_kernel<<<1, aux::kFilterThreads, 0, stream>>>(
        reinterpret_cast<const uint64_t*>(fwd), reinterpret_cast<const uint64_t*>(second), nbQueryDesc,
        reinterpret_cast<const uint64_t*>(bwd), nbTrainDesc, ratioNum, ratioDen, out, count);
    return check_launch("filter_matches");
}

extern "C" int rsfm_filter_matches_batch(const rsfm_filter_task* tasks, size_t nbTasks, uint32_t ratioNum,
                                         uint32_t ratioDen, hipStream_t stream)
{
    using namespace aux;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "filter batch: tasks is null");
    RSFM_RETURN_IF(ratioNum == 0 || ratioDen == 0, RSFM_ERR_INVALID, "filter batch: ratio must be num/den > 0");
    for (size_t b = 0; b < nbTasks; b += kFilterMaxTasks) {
        const size_t n = std::min<size_t>(kFilterMaxTasks, nbTasks - b);
        FilterBatch batch{};
        for (size_t i = 0; i < n; i++) {
            const rsfm_filter_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.count == nullptr, RSFM_ERR_INVALID, "filter batch: count is null");
            RSFM_RETURN_IF(t.nbQueryDesc > 0 && t.fwd == nullptr, RSFM_ERR_INVALID, "filter batch: fwd is null");
            RSFM_RETURN_IF(t.pts != nullptr && (t.kpQuery == nullptr || t.kpTrain == nullptr), RSFM_ERR_INVALID,
                           "filter batch: pts needs kpQuery and kpTrain");
            RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16 || reinterpret_cast<uintptr_t>(t.kpQuery) % 8 ||
                               reinterpret_cast<uintptr_t>(t.kpTrain) % 8,
                           RSFM_ERR_INVALID, "filter batch: pts must be 16-byte and keypoints 8-byte aligned");
            batch.t[i] = FilterTask{reinterpret_cast<const uint64_t*>(t.fwd), reinterpret_cast<const uint64_t*>(t.second),
                                    t.nbQueryDesc, reinterpret_cast<const uint64_t*>(t.bwd), t.nbTrainDesc, t.out,
                                    t.count, reinterpret_cast<const float2*>(t.kpQuery),
                                    reinterpret_cast<const float2*>(t.kpTrain), reinterpret_cast<float4*>(t.pts)};
        }
        filter_batch_kernel<<<(uint32_t)n, kFilterThreads, 0, stream>>>(batch, ratioNum, ratioDen);
        RSFM_CHECK(check_launch("filter_matches_batch"));
    }
    return RSFM_OK;
}
