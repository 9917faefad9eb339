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
// Vector-ALU matcher (the IDP path, bruteForceMatch.cu:560-620): u8 dot
// products with v_dot4_u32_u8, L2^2 = |a|^2 + |b|^2 - 2 a.b, exact in u32.
// A workgroup (256 threads) takes 128 query rows against a segment of 256
// train columns in four tiles of 64. Both operands sit in LDS K-major
// ([dword][row or column]); thread (tr = tid / 16, tc = tid % 16) owns rows
// 8 tr .. 8 tr + 7 and columns 4 tc .. 4 tc + 3 of a tile, so per K-dword two
// 16-byte reads of A and one of B feed 32 dot products.
//  * row side: key distance << 8 | column within the segment (u32, so ties go
//    to the lower column), a running minimum per row over the segment, then a
//    16-lane butterfly and one 64-bit atomicMin per row and segment;
//  * column side: key distance << 7 | row within the workgroup, minimum over
//    the thread's 8 rows, the wave's 4 row groups by lane swaps and the 4
//    waves through LDS, one 64-bit atomicMin per column and tile.
// Positions past the end load zeros and take the norm 2^23, so their
// distances are >= 2^23 > 128 * 255^2, the largest real one.
constexpr int kIdpThreads = 256;
constexpr int kIdpTM = 128;   // query rows per workgroup
constexpr int kIdpTN = 64;    // train columns per tile
constexpr int kIdpSeg = 256;  // train columns per workgroup
constexpr uint32_t kIdpPastNorm = 1u << 23;
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
    const uint32_t rowBase = blockIdx.x * kIdpTM;
    const uint32_t colBeg = blockIdx.y * kIdpSeg;
    if (rowBase >= T.nA || colBeg >= T.nB) return;
    const uint32_t nTiles = (min(colBeg + kIdpSeg, T.nB) - colBeg + kIdpTN - 1) / kIdpTN;
    __shared__ __attribute__((aligned(16))) uint32_t sA[32][kIdpTM];
    __shared__ __attribute__((aligned(16))) uint32_t sB[2][32][kIdpTN];
    __shared__ __attribute__((aligned(16))) uint32_t sNA[kIdpTM];
    __shared__ __attribute__((aligned(16))) uint32_t sNB[2][kIdpTN];
    __shared__ uint32_t sCol[2][4][kIdpTN];
    const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15, lane = tid & 63, wave = tid >> 6;

    // A: thread t stages 16 dwords of row t / 2 (the descriptors are only
    // 4-byte aligned, so dword loads)
    {
        const uint32_t r = (uint32_t)tid >> 1, row = rowBase + r;
        const bool valid = row < T.nA;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(T.A + (size_t)(valid ? row : 0u) * 128) + (tid & 1) * 16;
        uint32_t x[16], na = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) x[k] = valid ? src[k] : 0u;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            sA[(tid & 1) * 16 + k][r] = x[k];
            na = __builtin_amdgcn_udot4(x[k], x[k], na, false);
        }
        na += (uint32_t)__shfl_xor((int)na, 1);
        if ((tid & 1) == 0) sNA[r] = valid ? na : kIdpPastNorm;
    }
    // B tile of 64 columns: thread t loads 8 dwords of column t / 4
    uint32_t xb[8];
    auto load_b = [&](uint32_t tile) {
        const uint32_t col = colBeg + tile * kIdpTN + ((uint32_t)tid >> 2);
        const bool valid = col < T.nB;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(T.B + (size_t)(valid ? col : 0u) * 128) + (tid & 3) * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) xb[k] = valid ? src[k] : 0u;
    };
    auto store_b = [&](uint32_t tile, int buf) {
        const uint32_t c = (uint32_t)tid >> 2;
        uint32_t nb = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            sB[buf][(tid & 3) * 8 + k][c] = xb[k];
            nb = __builtin_amdgcn_udot4(xb[k], xb[k], nb, false);
        }
        nb += (uint32_t)__shfl_xor((int)nb, 1);
        nb += (uint32_t)__shfl_xor((int)nb, 2);
        if ((tid & 3) == 0) sNB[buf][c] = colBeg + tile * kIdpTN + c < T.nB ? nb : kIdpPastNorm;
    };
    load_b(0);
    store_b(0, 0);
    __syncthreads();

    uint32_t na[8];
#pragma unroll
    for (int i = 0; i < 8; i++) na[i] = sNA[tr * 8 + i];
    uint32_t rk[8];  // row tr * 8 + i: distance << 8 | column within the segment
#pragma unroll
    for (int i = 0; i < 8; i++) rk[i] = UINT32_MAX;

    for (uint32_t t = 0; t < nTiles; t++) {
        const int buf = (int)(t & 1);
        if (t + 1 < nTiles) load_b(t + 1);
        uint32_t acc[8][4];
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = 0;
#pragma unroll 4
        for (int k = 0; k < 32; k++) {
            const uint4 a0 = *reinterpret_cast<const uint4*>(&sA[k][tr * 8]);
            const uint4 a1 = *reinterpret_cast<const uint4*>(&sA[k][tr * 8 + 4]);
            const uint4 bv = *reinterpret_cast<const uint4*>(&sB[buf][k][tc * 4]);
            const uint32_t av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
            const uint32_t bw[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_udot4(av[i], bw[j], acc[i][j], false);
        }
        const uint4 nbv = *reinterpret_cast<const uint4*>(&sNB[buf][tc * 4]);
        const uint32_t nb[4] = {nbv.x, nbv.y, nbv.z, nbv.w};
        uint32_t ck[4] = {UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX};
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t d = na[i] + nb[j] - 2u * acc[i][j];
                rk[i] = min(rk[i], (d << 8) | (t * kIdpTN + tc * 4 + j));
                ck[j] = min(ck[j], (d << 7) | (uint32_t)(tr * 8 + i));
            }
        // columns: the wave's 4 row groups (lanes 16 apart), then the 4 waves
#pragma unroll
        for (int j = 0; j < 4; j++) {
            ck[j] = min(ck[j], (uint32_t)__shfl_xor((int)ck[j], 16));
            ck[j] = min(ck[j], (uint32_t)__shfl_xor((int)ck[j], 32));
        }
        if (lane < 16) {
#pragma unroll
            for (int j = 0; j < 4; j++) sCol[buf][wave][tc * 4 + j] = ck[j];
        }
        if (t + 1 < nTiles) store_b(t + 1, buf ^ 1);
        __syncthreads();
        if (T.bestB != nullptr && tid < kIdpTN) {
            const uint32_t m = min(min(sCol[buf][0][tid], sCol[buf][1][tid]), min(sCol[buf][2][tid], sCol[buf][3][tid]));
            const uint32_t col = colBeg + t * kIdpTN + (uint32_t)tid;
            if (col < T.nB && m < (1u << 30))
                atomicMin(reinterpret_cast<unsigned long long*>(T.bestB + col),
                          (unsigned long long)make_key(m >> 7, rowBase + (m & 127)));
        }
    }
    // rows: minimum over the 16 column groups, then lane tc < 8 takes row tr * 8 + tc
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) rk[i] = min(rk[i], (uint32_t)__shfl_xor((int)rk[i], o));
    uint32_t mine = rk[0];
#pragma unroll
    for (int i = 1; i < 8; i++) mine = tc == i ? rk[i] : mine;
    const uint32_t row = rowBase + tr * 8 + tc;
    if (tc < 8 && row < T.nA && mine < (1u << 31))
        atomicMin(reinterpret_cast<unsigned long long*>(T.bestA + row),
                  (unsigned long long)make_key(mine >> 8, colBeg + (mine & 255)));
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
        idp_kernel<<<dim3(div_up(maxA, kIdpTM), div_up(maxB, kIdpSeg), (uint32_t)n), kIdpThreads, 0, stream>>>(batch);
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
