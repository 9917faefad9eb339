// bf_match.hip - brute-force SIFT matching on gfx950 int8 matrix cores.
//
// Contract (reference: kernelBruteForceMatchIMMA, bruteForceMatchIMMA.cu:551-586,
// packed keys bruteForceMatch.cuh:108-261): for every query i the lexicographic
// minimum of (L2^2(i,j), j) over all train j, and for every train j the minimum
// of (L2^2(i,j), i) over all queries i. Distances are exact integers.
//
// Design (own, for CDNA4):
//  * u8 descriptors are mapped to signed bytes u = 127 - a (rows) and
//    w = b - 128 (columns) so one v_mfma_i32_16x16x64_i8 chain gives
//    dot = u.w and  L2^2 = NA' + NB' + 2 dot  with NA' = |a-128|^2 and
//    NB' = |b-127|^2 - 128 (both in [-128, 2^21]); 4-bit descriptors are
//    widened to bytes as they are loaded.
//  * Every call first writes the train side in its staging form (w and
//    NB' << 8 | tag per column, padded to whole 64-column chunks) into a
//    workspace (the caller's for top-2, a stream-ordered allocation for
//    top-1); the tile copies it into an LDS ring with direct global-to-LDS
//    loads, so staging costs neither registers nor VALU in the main loop.
//  * A workgroup is 8 waves x 32 query rows, their operands resident in
//    registers, over a segment of at most 4096 train columns; two
//    workgroups share a CU.
//  * The MFMA chain starts from H_i = NA'_i >> 1 instead of zero, so the
//    single key R = (acc << 9) + (NB'_j << 8 | tag) = 256 (d - p_i) + tag
//    (p_i = NA'_i & 1) orders each row's columns exactly and each column's
//    rows up to the parity p_i. One v_lshl_add per product; v_min3 folds
//    the row best (v_med3 + v_min3 the row top-2) and the column minimum.
//  * Train side: per column the workgroup keeps its smallest d - p (and the
//    first wave holding it) in LDS and lowers a global per-column minimum
//    by atomicMin; at the end of the tile every column within 1 of that
//    minimum is recomputed exactly on MFMA and merged by a 64-bit atomicMin.
//  * Segments split the train range when one launch has too few row blocks
//    to fill 256 CUs; row results then merge by 64-bit atomicMin (top-1) or
//    through a small workspace and a merge kernel (top-2).
#include "internal.hpp"
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace rsfm_amd {
namespace bf {

constexpr int32_t kNone = INT32_MAX;

// Prepared column offset of the columns past nB, whose staged operands
// are zero: with the accumulators started at H_i = NA'_i >> 1 such a column's
// key is 256 (NA'_i - p_i) + kPastCols + tag. Over all u8 b,
// |a - b|^2 - NA' <= 128 * 48896 (at a byte of 255), so kPastCols above
// 256 * 128 * 48896 + 255 keeps that key above every real column's key
// 256 (d - p_i) + tag, and NA' <= 128 * 128^2 keeps it below 2^31.
constexpr int32_t kPastCols = 0x5FC00000;
static_assert(256LL * 128 * 48896 + 255 < kPastCols, "past-nB key must exceed every real key");
static_assert(256LL * 128 * 16384 + kPastCols + 255 <= INT32_MAX, "past-nB key must not wrap");

// 16 bytes of the (expanded) descriptor at byte offset off of the 128-value row
template <int kBits>
__device__ inline v4i load16(const uint8_t* desc, int off)
{
    if constexpr (kBits == 8) {
        return *reinterpret_cast<const v4i*>(desc + off);
    } else {
        // 4-bit: value 2b is the low nibble of byte b, value 2b+1 the high one
        const uint2 x = *reinterpret_cast<const uint2*>(desc + off / 2);
        v4i r;
        const uint32_t w[2] = {x.x, x.y};
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t lo = w[h] & 0x0F0F0F0Fu;
            const uint32_t hn = (w[h] >> 4) & 0x0F0F0F0Fu;
            r[2 * h + 0] = (int)__builtin_amdgcn_perm(hn, lo, 0x05010400u);
            r[2 * h + 1] = (int)__builtin_amdgcn_perm(hn, lo, 0x07030602u);
        }
        return r;
    }
}

__device__ inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
__device__ inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
__device__ inline uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ inline uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

// (a << 9) + b: one v_lshl_add_u32
__device__ inline int32_t shl9_add(int32_t a, int32_t b) { return (int32_t)(((uint32_t)a << 9) + (uint32_t)b); }
// middle of three: one v_med3_i32
__device__ inline int32_t med3(int32_t a, int32_t b, int32_t c) { return imax(imin(a, b), imin(imax(a, b), c)); }
// Fold two new keys x0, x1 into a running top-2 (m1 < m2). The smallest of
// {m1, x0, x1} is the new m1; the second smallest of {m1, m2, x0, x1} is
// min(m2, med3(m1, x0, x1)) because m2 >= m1. Three instructions per pair.
__device__ inline void top2_fold2(int32_t& m1, int32_t& m2, int32_t x0, int32_t x1)
{
    const int32_t t = med3(m1, x0, x1);
    // asm: the compiler would reuse med3's inner min(m1, x0) and spend two v_min
    asm("v_min3_i32 %0, %0, %1, %2" : "+v"(m1) : "v"(x0), "v"(x1));
    m2 = imin(m2, t);
}
// DPP lane exchanges inside groups of 8 consecutive lanes
__device__ inline int32_t dpp_xor1(int32_t x) { return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false); }
__device__ inline int32_t dpp_xor2(int32_t x) { return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false); }
// lane i <-> lane 7 - i of each group of 8 (row_half_mirror)
__device__ inline int32_t dpp_mirror8(int32_t x) { return __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false); }

// 16-B slot swizzle of a staged column: columns c and c^1 sit in different
// 128-B halves of a 256-B bank row, and the eight columns of one parity map
// slot s to eight different slots, so both 8- and 16-lane b128 phases are
// bank-conflict free
__device__ inline int swz(int c) { return ((c >> 1) & 7) ^ ((c & 1) << 2); }

// Direct global -> LDS copies (gfx950 global_load_lds): every lane moves
// `size` bytes from its own global address to lds + size * lane; lds must be
// the same for the whole wave. Completion is counted by vmcnt.
__device__ inline void dma16(const void* src, void* lds)
{
    __builtin_amdgcn_global_load_lds(
        reinterpret_cast<const __attribute__((address_space(1))) void*>(reinterpret_cast<uintptr_t>(src)),
        reinterpret_cast<__attribute__((address_space(3))) void*>(reinterpret_cast<uintptr_t>(lds)), 16, 0, 0);
}
__device__ inline void dma4(const void* src, void* lds)
{
    __builtin_amdgcn_global_load_lds(
        reinterpret_cast<const __attribute__((address_space(1))) void*>(reinterpret_cast<uintptr_t>(src)),
        reinterpret_cast<__attribute__((address_space(3))) void*>(reinterpret_cast<uintptr_t>(lds)), 4, 0, 0);
}
// a wave-uniform address as an opaque scalar value: a per-lane offset added to
// it is not reassociated into a per-lane pointer advanced every trip, so a
// copy's address costs one 64-bit vector add (the chunk step stays scalar)
__device__ inline const uint8_t* uniform_ptr(const void* p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<const uint8_t*>(((uint64_t)hi << 32) | lo);
}
// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt left at their maxima)
__device__ inline void wait_vm0() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Matcher tile over the prepared train form (rsfm_bf_match_top2 and the
// batched pipeline with kTop2; rsfm_bf_match_u8 / _u4 without: the row side
// then keeps only its best), on v_mfma_i32_16x16x64_i8. One workgroup = 8 waves x 32 query
// rows, two workgroups (four waves per SIMD) per CU. Lane l of a wave works
// on column class l % 16 and row group q = l / 16:
//  * query operands stay in registers: for row block rb (16 rows) and K-step
//    ks (64 bytes) the lane holds bytes [64 ks + 16 q, +16) of row
//    rb * 16 + l % 16 as u = 127 - a (rows past nA repeat row nA - 1, so they
//    tie with it and never beat it on the train side);
//  * the segment streams through a 4-slot LDS ring in 64-column chunks copied
//    by direct global-to-LDS loads of the prepared form (w = b - 128 and
//    NB' << 8 | tag per column, written once per call by prep_train_kernel);
//    a chunk is 4 column blocks of 16, so a chunk's 16 MFMAs (2 row blocks x
//    4 column blocks x 2 K-steps) give the lane 32 products: rows
//    rb * 16 + 4 q + j (j < 4) of columns cb * 16 + l % 16; two chunks per
//    trip between barriers, the upper four waves running the keys of a
//    trip's second chunk half a trip late (on each SIMD one wave issues
//    MFMAs while the other runs keys);
//  * the MFMA chains start at H_i = NA'_i >> 1, so the single key
//    R = (acc << 9) + (NB'_j << 8 | tag) = 256 (d - p_i) + tag
//    (p_i = NA'_i & 1, tag = 4 chunk + cb within the segment, hence at most
//    4096 columns per segment) orders each row's columns exactly and each
//    column's rows up to p_i: one v_lshl_add per product, v_med3 + v_min3 for
//    the row top-2 per (row, class), v_min3 for the column minimum, and two
//    lane swaps (v_permlane32_swap, v_permlane16_swap) that leave lane L with
//    column L's minimum over the wave's rows;
//  * train side: per column the block's smallest R with the first wave
//    holding it (in LDS), a global per-column minimum by atomicMin, and at
//    the end of the tile the exact distances of the candidate (column, wave)
//    pairs on MFMA, merged into bestB by atomicMin.
template <int kBits, bool kTop2>
__device__ __forceinline__ void match_tile_top2(const DevTask& T, uint32_t rowBase, uint32_t colBeg, uint32_t seg)
{
    constexpr int kWaves = 8;
    constexpr int kChunkBytes = kBN * 128;
    constexpr int kRing = 4;  // ring slot of chunk c: c % 4
    constexpr int kPSlots = 8;
    constexpr int kColSlots = 8;
    constexpr int kCS = 72;
    constexpr int kTailStride = 17;
    constexpr int kUnionBytes = kRing * kChunkBytes;
    static_assert(kBM * kTailStride * 4 <= kUnionBytes, "row transpose exceeds the ring");

    const uint32_t colEnd = min(colBeg + T.segCols, T.nB);
    const uint32_t nChunks = (colEnd - colBeg + kBN - 1) / kBN;

    __shared__ __attribute__((aligned(16))) uint8_t sU[kUnionBytes];
    __shared__ int32_t sP[kPSlots][kBN];
    __shared__ int32_t sCol[kColSlots][kWaves][kCS];
    __shared__ __attribute__((aligned(16))) int32_t sQ[kBM];  // NA'
    __shared__ int32_t sPart[kMaxSegColsTop2];                 // this block's train-side entries
    __shared__ uint32_t sNres;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c16 = lane & 15, q = lane >> 4;

    // ---- query operands and NA' = |a - 128|^2 (four lanes per row)
    v4i aop[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        const uint32_t row = min(rowBase + wave * 32 + rb * 16 + c16, T.nA - 1);
        int32_t nap = 0;
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const v4i x = load16<kBits>(T.A + (size_t)row * (kBits * 16), ks * 64 + q * 16);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int xa = x[k] ^ (int)0x80808080u;
                nap = __builtin_amdgcn_sdot4(xa, xa, nap, false);
                aop[rb][ks][k] = x[k] ^ 0x7F7F7F7F;
            }
        }
        nap += __shfl_xor(nap, 16);
        nap += __shfl_xor(nap, 32);
        if (q == 0) sQ[wave * 32 + rb * 16 + c16] = nap;
    }

    int32_t m1[2][4], m2[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            m1[rb][j] = kNone;
            m2[rb][j] = kNone;
        }
    v4i Hv[2];  // accumulator start H = NA' >> 1 of rows rb * 16 + 4 q + j

    // lane L of wave w fills column w*8 + L/8, slot L%8 of the chunk's ring
    // slot, i.e. it reads source slot (L%8) ^ swz(column); wave 0 also copies
    // the chunk's 64 column offsets
    const uint8_t* const segW = T.W + (size_t)colBeg * 128;
    const int32_t* const segNB = T.NB + colBeg;
    const uint32_t dmaOff = (uint32_t)((wave * 8 + (lane >> 3)) * 128 + (((lane & 7) ^ swz(wave * 8 + (lane >> 3))) << 4));
    auto stage_dma = [&](uint32_t c) {
        const uint8_t* chunkW = uniform_ptr(segW + (size_t)c * (kBN * 128));
        dma16(chunkW + dmaOff, sU + (c % kRing) * kChunkBytes + wave * 1024);
        if (wave == 0) dma4(uniform_ptr(segNB + (size_t)c * kBN) + lane * 4, &sP[c % kPSlots][0]);
    };
    // wave k % 8 reduces chunk k's 8 wave minima per column (one lane per
    // column): E = Rmin >> 8 to T.colMin by atomicMin, and to the block's
    // entry in sPart with the tag replaced by the first wave holding Rmin
    // (bits 0-2) and bit 3 set when another wave comes within one unit
    // (E_w <= E + 1)
    auto flush = [&](uint32_t k) {
        if (wave != (int)(k & 7)) return;
        const int32_t* src = &sCol[k % kColSlots][0][lane];
        int32_t k1 = kNone, k2 = kNone;
#pragma unroll
        for (int w = 0; w < kWaves; w += 2)
            top2_fold2(k1, k2, (src[w * kCS] & ~0xFF) | w, (src[(w + 1) * kCS] & ~0xFF) | (w + 1));
        const uint32_t gcol = colBeg + k * kBN + lane;
        if (gcol < colEnd) {
            sPart[k * kBN + lane] = k1 | ((k2 >> 8) <= (k1 >> 8) + 1 ? 8 : 0);
            atomicMin(T.colMin + gcol, k1 >> 8);
        }
    };
    // lane L ends with the minimum over the wave's rows of column L, from the
    // lanes' per-class minima x[cb] (rows of group q, column cb * 16 + l % 16)
    auto col_reduce = [&](const int32_t (&x)[4]) -> int32_t {
        const auto a = __builtin_amdgcn_permlane32_swap((uint32_t)x[0], (uint32_t)x[2], false, false);
        const auto b = __builtin_amdgcn_permlane32_swap((uint32_t)x[1], (uint32_t)x[3], false, false);
        const int32_t t02 = imin((int32_t)a[0], (int32_t)a[1]);
        const int32_t t13 = imin((int32_t)b[0], (int32_t)b[1]);
        const auto s = __builtin_amdgcn_permlane16_swap((uint32_t)t02, (uint32_t)t13, false, false);
        return imin((int32_t)s[0], (int32_t)s[1]);
    };
    const bool doCols = T.bestB != nullptr;

    // chunk c's 16 MFMAs (operand reads of one K-step at a time)
    auto mfma_chunk = [&](uint32_t c, v4i (&acc)[2][4]) {
        const uint8_t* sb = sU + (c % kRing) * kChunkBytes;
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            v4i b[4];
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                const int col = cb * 16 + c16;
                b[cb] = *reinterpret_cast<const v4i*>(sb + col * 128 + (((4 * ks + q) ^ swz(col)) << 4));
            }
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++)
                    acc[rb][cb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aop[rb][ks], b[cb], ks == 0 ? Hv[rb] : acc[rb][cb], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // chunk c's keys: row top-2 folds, column minima to sCol
    auto keys_chunk = [&](uint32_t c, const v4i (&acc)[2][4]) {
        int32_t P[4];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) P[cb] = sP[c % kPSlots][cb * 16 + c16];
        int32_t cm[4];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int jp = 0; jp < 4; jp += 2) {
                int32_t r[2][4];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int j = jp + h;
#pragma unroll
                    for (int cb = 0; cb < 4; cb++) r[h][cb] = shl9_add(acc[rb][cb][j], P[cb]);
                    if constexpr (kTop2) {
                        const int32_t t1 = med3(m1[rb][j], r[h][0], r[h][1]);
                        asm("v_min3_i32 %0, %0, %1, %2" : "+v"(m1[rb][j]) : "v"(r[h][0]), "v"(r[h][1]));
                        const int32_t t2 = med3(m1[rb][j], r[h][2], r[h][3]);
                        asm("v_min3_i32 %0, %0, %1, %2" : "+v"(m1[rb][j]) : "v"(r[h][2]), "v"(r[h][3]));
                        asm("v_min3_i32 %0, %0, %1, %2" : "+v"(m2[rb][j]) : "v"(t1), "v"(t2));
                    } else {
                        asm("v_min3_i32 %0, %0, %1, %2" : "+v"(m1[rb][j]) : "v"(r[h][0]), "v"(r[h][1]));
                        asm("v_min3_i32 %0, %0, %1, %2" : "+v"(m1[rb][j]) : "v"(r[h][2]), "v"(r[h][3]));
                    }
                }
#pragma unroll
                for (int cb = 0; cb < 4; cb++) {
                    if (rb == 0 && jp == 0)
                        cm[cb] = imin(r[0][cb], r[1][cb]);
                    else
                        asm("v_min3_i32 %0, %0, %1, %2" : "+v"(cm[cb]) : "v"(r[0][cb]), "v"(r[1][cb]));
                }
            }
        sCol[c % kColSlots][wave][lane] = col_reduce(cm);
    };

    stage_dma(0);
    if (nChunks > 1) stage_dma(1);
    wait_vm0();
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        const v4i q4 = *reinterpret_cast<const v4i*>(sQ + wave * 32 + rb * 16 + 4 * q);
#pragma unroll
        for (int j = 0; j < 4; j++) Hv[rb][j] = q4[j] >> 1;
    }

    // Two chunks per trip, one barrier per trip. The waves of the upper half
    // (the second wave on each SIMD) run the keys of a trip's second chunk at
    // the start of the next trip, so on every SIMD one wave issues MFMAs while
    // the other runs key epilogues. A trip's first chunk copies chunks c+2 and
    // c+3 (their slots were last read before the previous barrier); a chunk's
    // column partials are complete one trip after its own and flushed then.
    const bool lagging = wave >= kWaves / 2;
    v4i accA[2][4], accB[2][4];
    uint32_t c = 0;
    for (; c + 2 <= nChunks; c += 2) {
        if (lagging && c >= 2) keys_chunk(c - 1, accB);
        if (c + 2 < nChunks) stage_dma(c + 2);
        if (c + 3 < nChunks) stage_dma(c + 3);
        if (doCols && c >= 2) flush(c - 2);
        if (doCols && c >= 4) flush(c - 3);
        mfma_chunk(c, accA);
        keys_chunk(c, accA);
        mfma_chunk(c + 1, accB);
        if (!lagging) keys_chunk(c + 1, accB);
        wait_vm0();
        __syncthreads();
    }
    const uint32_t firstUnflushed = c >= 3 ? c - 3 : 0;
    if (lagging && c >= 2) keys_chunk(c - 1, accB);
    if (c < nChunks) {
        // the odd last chunk: its copy was issued by the last trip (or before the loop)
        mfma_chunk(c, accA);
        keys_chunk(c, accA);
    }
    __syncthreads();
    if (doCols)
        for (uint32_t k = firstUnflushed; k < nChunks; k++) flush(k);

    // ---- row results: transpose the per-(row, class) bests through LDS
    int32_t* L = reinterpret_cast<int32_t*>(sU);
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) L[(wave * 32 + rb * 16 + 4 * q + j) * kTailStride + c16] = m1[rb][j];
    __syncthreads();
    int32_t b1 = kNone, b2 = kNone;
    int c1 = 0, c2 = 0;
    if (tid < kBM) {
        for (int cc = 0; cc < 16; cc++) {
            const int32_t x = L[tid * kTailStride + cc];
            if (x < b1) {
                b2 = b1; c2 = c1;
                b1 = x; c1 = cc;
            } else if (x < b2) {
                b2 = x; c2 = cc;
            }
        }
    }
    if constexpr (kTop2) {
        __syncthreads();
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int j = 0; j < 4; j++) L[(wave * 32 + rb * 16 + 4 * q + j) * kTailStride + c16] = m2[rb][j];
        __syncthreads();
    }
    if (tid < kBM) {
        if constexpr (kTop2) {
            const int32_t x = L[tid * kTailStride + c1];
            if (x < b2 || (x == b2 && c1 < c2)) {
                b2 = x;
                c2 = c1;
            }
        }
        const uint32_t row = rowBase + tid;
        if (row < T.nA) {
            // keys 256 (d - p) + tag; keys of columns past nB start at
            // 256 (NA' - p) + kPastCols, above every real column's key
            const int32_t p = sQ[tid] & 1;
            const int32_t past = (sQ[tid] - p) * 256 + kPastCols;
            auto to_key = [&](int32_t m, int cls) -> uint64_t {
                if (m == kNone || m >= past) return kNoMatch;
                return make_key((uint32_t)((m >> 8) + p), colBeg + (uint32_t)(m & 0xFF) * 16u + (uint32_t)cls);
            };
            const uint64_t k1 = to_key(b1, c1);
            if constexpr (kTop2) {
                const uint64_t k2 = to_key(b2, c2);
                if (T.nSeg == 1) {
                    T.bestA[row] = k1;
                    T.secondA[row] = k2;
                } else {
                    uint64_t* pp = T.part + ((size_t)seg * T.nA + row) * 2;
                    pp[0] = k1;
                    pp[1] = k2;
                }
            } else if (T.nSeg == 1) {
                T.bestA[row] = k1;
            } else if (k1 != kNoMatch) {
                // several segments: merged in place (results start at "no match")
                atomicMin(reinterpret_cast<unsigned long long*>(T.bestA + row), (unsigned long long)k1);
            }
        }
    }

    // ---- train side. For column j this block's entry (sPart) holds E = min
    // over its rows of (d - p_i); e = min over all blocks (T.colMin). The lowest row of
    // smallest d has d - p <= e + 1 (a row with E = e has d <= e + 1), so
    // only blocks with E <= e + 1 hold it, and in such a block only the first
    // wave named in the entry (all 8 when its bit 3 is set). T.colMin still
    // decreases while other blocks run, so this block takes every column
    // with E <= colMin + 1 as read now: a superset of its true share. Each
    // listed (column, wave) gets its exact (d, row) minimum over the wave's
    // 32 rows on MFMA, merged into bestB by a 64-bit atomicMin.
    if (!doCols) return;
    constexpr uint32_t kBuf = 128, kPer = kMaxSegColsTop2 / (kWaves * 64);
    uint16_t* sListJ = reinterpret_cast<uint16_t*>(sU);
    uint8_t* sListM = reinterpret_cast<uint8_t*>(sListJ + kMaxSegColsTop2);
    uint32_t* sBuf = reinterpret_cast<uint32_t*>(sListM + kMaxSegColsTop2);  // [kWaves][kBuf]
    static_assert(kMaxSegColsTop2 * 3 + kWaves * kBuf * 4 <= kUnionBytes, "resolution lists exceed the ring");
    uint32_t* buf = sBuf + wave * kBuf;
    // one batch of up to 64 listed columns of this wave: lane L ends with
    // column L's d << 8 | local row = (u.w << 9) + (NA' << 8 | local row) +
    // (NB' << 8), minimised over the wave's rows
    auto resolve_batch = [&](uint32_t cnt) {
        int32_t cm[4];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            const uint32_t jc = buf[min((uint32_t)(cb * 16 + c16), cnt - 1)];
            v4i acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const v4i b = *reinterpret_cast<const v4i*>(T.W + (size_t)jc * 128 + (4 * ks + q) * 16);
#pragma unroll
                for (int rb = 0; rb < 2; rb++) acc[rb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aop[rb][ks], b, acc[rb], 0, 0, 0);
            }
            int32_t k = kNone;
#pragma unroll
            for (int rb = 0; rb < 2; rb++) {
                const v4i q4 = *reinterpret_cast<const v4i*>(sQ + wave * 32 + rb * 16 + 4 * q);
#pragma unroll
                for (int j = 0; j < 4; j++)
                    k = imin(k, shl9_add(acc[rb][j], q4[j] * 256 + wave * 32 + rb * 16 + 4 * q + j));
            }
            cm[cb] = k;
        }
        const int32_t key = col_reduce(cm);
        if ((uint32_t)lane < cnt) {
            const uint32_t j = buf[lane];
            const uint32_t dk = (uint32_t)(key + (T.NB[j] & ~0xFF));
            atomicMin(reinterpret_cast<unsigned long long*>(T.bestB + j),
                      (unsigned long long)make_key(dk >> 8, rowBase + (dk & 0xFF)));
        }
    };
    if (tid == 0) sNres = 0;
    // every load of the scan in flight at once
    int32_t x[kPer], e[kPer];
#pragma unroll
    for (uint32_t t = 0; t < kPer; t++) {
        const uint32_t j = colBeg + t * (kWaves * 64) + (uint32_t)tid;
        x[t] = j < colEnd ? sPart[j - colBeg] : INT32_MAX;
        e[t] = j < colEnd ? T.colMin[j] : 0;
    }
    __syncthreads();  // the row results' reads of sU, and sNres
#pragma unroll
    for (uint32_t t = 0; t < kPer; t++) {
        const bool cand = (x[t] >> 8) <= e[t] + 1;
        const uint64_t m = __ballot(cand);
        if (m == 0) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&sNres, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, 0);
        if (cand) {
            const uint32_t pos = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
            sListJ[pos] = (uint16_t)(t * (kWaves * 64) + (uint32_t)tid);
            sListM[pos] = (uint8_t)((x[t] & 8) ? 0xFFu : (1u << (x[t] & 7)));
        }
    }
    __syncthreads();
    const uint32_t n = sNres;
    // this wave's entries, gathered 64 at a time into its buffer
    uint32_t pend = 0;
    for (uint32_t e0 = 0; e0 < n; e0 += 64) {
        const uint32_t ei = e0 + (uint32_t)lane;
        const bool sel = ei < n && ((sListM[ei] >> wave) & 1);
        const uint64_t m = __ballot(sel);
        if (sel) buf[pend + __popcll(m & ((1ull << lane) - 1))] = colBeg + sListJ[ei];
        pend += (uint32_t)__popcll(m);
        if (pend >= 64) {
            resolve_batch(64);
            // move the rest (at most 63) to the front
            const uint32_t rest = pend - 64;
            const uint32_t v = (uint32_t)lane < rest ? buf[64 + lane] : 0u;
            if ((uint32_t)lane < rest) buf[lane] = v;
            pend = rest;
        }
    }
    if (pend > 0) resolve_batch(pend);
}

// grid launch: x = row block, y = segment, z = task (XCD-local order).
// __launch_bounds__(512, 4): 512 threads, at least four waves per SIMD, i.e.
// at most 128 VGPRs, so two workgroups fit a CU.
// Top-1 (rsfm_bf_match_u8 / _u4, 8- or 4-bit descriptors).
template <int kBits>
__global__ __launch_bounds__(512, 4) void match_top1_kernel(const TaskBatch batch)
{
    const Wg3 wg = xcd_local_block();
    const DevTask& T = batch.t[wg.z];
    const uint32_t rowBase = wg.x * kBM;
    const uint32_t colBeg = wg.y * T.segCols;
    if (rowBase >= T.nA || colBeg >= T.nB) return;
    match_tile_top2<kBits, false>(T, rowBase, colBeg, wg.y);
}

__global__ __launch_bounds__(512, 4) void match_top2_kernel(const TaskBatch batch)
{
    const Wg3 wg = xcd_local_block();
    const DevTask& T = batch.t[wg.z];
    const uint32_t rowBase = wg.x * kBM;
    const uint32_t colBeg = wg.y * T.segCols;
    if (rowBase >= T.nA || colBeg >= T.nB) return;
    match_tile_top2<8, true>(T, rowBase, colBeg, wg.y);
}

// Persistent launch over a tile table: the tiles are split into kNumXcd lists
// (xcdOff[x] .. xcdOff[x+1]), one per XCD; workgroup b serves list b % kNumXcd
// (the XCD the hardware dispatches it to), striding by the workgroups of that
// XCD, so all tiles of one task share one L2. gridDim.x is a multiple of
// kNumXcd.
__global__ __launch_bounds__(512, 4) void match_tiles_kernel(const DevTask* __restrict__ tasks,
                                                          const TileDesc* __restrict__ tiles,
                                                          const uint32_t* __restrict__ xcdOff)
{
    const uint32_t xcd = blockIdx.x % kNumXcd, per = gridDim.x / kNumXcd;
    const uint32_t end = xcdOff[xcd + 1];
    for (uint32_t i = xcdOff[xcd] + blockIdx.x / kNumXcd; i < end; i += per) {
        const TileDesc d = tiles[i];
        const DevTask T = tasks[d.task];
        __syncthreads();  // the previous tile's last LDS reads
        match_tile_top2<8, true>(T, d.rowBase, d.colBeg, d.seg);
    }
}

// The same tile table, one workgroup per tile: workgroup b takes element b / 8
// of XCD list b % 8 (the XCD the hardware dispatches it to). The body is
// the tile alone, without the persistent loop's extra live registers.
__global__ __launch_bounds__(512, 4) void match_tile_grid_kernel(const DevTask* __restrict__ tasks,
                                                              const TileDesc* __restrict__ tiles,
                                                              const uint32_t* __restrict__ xcdOff)
{
    const uint32_t xcd = blockIdx.x % kNumXcd;
    const uint32_t i = xcdOff[xcd] + blockIdx.x / kNumXcd;
    if (i >= xcdOff[xcd + 1]) return;
    const TileDesc d = tiles[i];
    const DevTask T = tasks[d.task];
    match_tile_top2<8, true>(T, d.rowBase, d.colBeg, d.seg);
}

// staging form of B (DevTask::W, NB): eight lanes per column, one 16-B slot
// each (4-bit descriptors widened to bytes); a workgroup of 256 covers 32
// columns and the padded column count is a multiple of 64, so a workgroup is
// wholly inside or outside it. The same threads also set results to "no
// match": kInitNone leaves them (the pipeline clears its own), kInitTop2 sets
// the train-side results (merged by atomicMin) and, for a task without train
// descriptors, the query-side ones (no tile writes them), kInitTop1 sets the
// train-side and all query-side results (several segments merge them by
// atomicMin).
enum : int { kInitNone = 0, kInitTop2 = 1, kInitTop1 = 2 };
template <int kBits, class Src>
__global__ __launch_bounds__(256) void prep_train_kernel(const Src batch, int init)
{
    const DevTask& T = batch[blockIdx.z];
    const uint32_t nPad = (T.nB + kBN - 1) / kBN * kBN;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (init != kInitNone) {
        if (T.bestB != nullptr && t < T.nB) T.bestB[t] = kNoMatch;
        if ((T.nB == 0 || init == kInitTop1) && t < T.nA) {
            T.bestA[t] = kNoMatch;
            if (T.secondA != nullptr) T.secondA[t] = kNoMatch;
        }
    }
    const uint32_t col = t >> 3, sp = t & 7;
    if (T.W == nullptr || col >= nPad) return;
    const bool valid = col < T.nB;
    v4i x = {0, 0, 0, 0};
    if (valid) x = load16<kBits>(T.B + (size_t)col * (kBits * 16), sp * 16);
    int32_t acc = 0;
    v4i w;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        w[q] = x[q] ^ (int)0x80808080u;
        if (!valid) w[q] = 0;
        acc = __builtin_amdgcn_sdot4(w[q], w[q], acc, false);
        acc = __builtin_amdgcn_sdot4(w[q], 0x02020202, acc, false);
    }
    *reinterpret_cast<v4i*>(T.W + (size_t)col * 128 + sp * 16) = w;
    acc += dpp_xor1(acc);
    acc += dpp_xor2(acc);
    acc += dpp_mirror8(acc);
    if (sp == 0) {
        // 16-column block within the segment, the low byte of a row key
        const int32_t tag = (int32_t)((col % T.segCols) >> 4);
        T.NB[col] = (valid ? acc * 256 : kPastCols) | tag;
        if (valid) T.colMin[col] = INT32_MAX;
    }
}

size_t prep_bytes(uint32_t nA, uint32_t nB)
{
    if (nA == 0 || nB == 0) return 0;
    const size_t nPad = div_up(nB, (uint32_t)kBN) * (size_t)kBN;
    return align_up(nPad * 128, 256) + align_up(nPad * 4, 256) + align_up((size_t)nB * 4, 256);
}

void place_prep(DevTask& t, uint8_t* ws)
{
    if (t.nA == 0 || t.nB == 0) return;
    const size_t nPad = div_up(t.nB, (uint32_t)kBN) * (size_t)kBN;
    t.W = ws;
    ws += align_up(nPad * 128, 256);
    t.NB = reinterpret_cast<int32_t*>(ws);
    ws += align_up(nPad * 4, 256);
    t.colMin = reinterpret_cast<int32_t*>(ws);
}

int launch_prep_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxB, hipStream_t stream)
{
    if (nTasks == 0 || maxB == 0) return RSFM_OK;
    const uint32_t nPad = div_up(maxB, (uint32_t)kBN) * kBN;
    prep_train_kernel<8><<<dim3(nPad * 8 / 256, 1, nTasks), 256, 0, stream>>>(TaskTable{tasks}, kInitNone);
    return check_launch("bf prep table");
}



template <class Src>
__global__ void merge_top2_kernel(const Src batch)
{
    const DevTask& T = batch[blockIdx.z];
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= T.nA || T.nSeg <= 1) return;
    uint64_t k1 = kNoMatch, k2 = kNoMatch;
    // segments at or past this task's train count were never written
    const uint32_t nUsed = min(T.nSeg, (T.nB + T.segCols - 1) / T.segCols);
    for (uint32_t s = 0; s < nUsed; s++) {
        const uint64_t* p = T.part + ((size_t)s * T.nA + row) * 2;
        const uint64_t o1 = p[0], o2 = p[1];
        k2 = umin64(umax64(k1, o1), umin64(k2, o2));
        k1 = umin64(k1, o1);
    }
    T.bestA[row] = k1;
    T.secondA[row] = k2;
}

// Launch geometry shared by the size query and the launch.
struct Plan {
    uint32_t gridX = 0, nSeg = 1, segCols = kBN;
    uint32_t maxA = 0, maxB = 0;
};

static Plan make_plan(const DevTask* t, size_t n, uint32_t maxSegCols)
{
    Plan p;
    for (size_t i = 0; i < n; i++) {
        p.maxA = std::max(p.maxA, t[i].nA);
        p.maxB = std::max(p.maxB, t[i].nB);
    }
    p.gridX = div_up(p.maxA, kBM);
    if (p.maxA == 0 || p.maxB == 0) return p;
    // aim for >= 512 workgroups (2 per CU) but keep >= 1024 columns per segment
    const uint32_t wg = p.gridX * (uint32_t)n;
    uint32_t nSeg = div_up(512, wg);
    nSeg = std::min(nSeg, std::max(1u, p.maxB / 1024));
    nSeg = std::max(nSeg, div_up(p.maxB, maxSegCols));
    uint32_t seg = div_up(p.maxB, nSeg);
    seg = div_up(seg, kBN) * kBN;
    p.segCols = std::min(seg, maxSegCols);
    p.nSeg = div_up(p.maxB, p.segCols);
    return p;
}

static int validate_task(const rsfm_bf_task& t, int bits)
{
    const size_t align = bits == 8 ? 16 : 8;
    RSFM_RETURN_IF(t.nbDescA > 0 && (t.descA == nullptr || t.bestMatchA == nullptr), RSFM_ERR_INVALID,
                   "bf task: descA/bestMatchA must be set when nbDescA > 0");
    RSFM_RETURN_IF(t.nbDescB > 0 && t.descB == nullptr, RSFM_ERR_INVALID, "bf task: descB must be set when nbDescB > 0");
    RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(t.descA) | reinterpret_cast<uintptr_t>(t.descB)) % align != 0,
                   RSFM_ERR_INVALID, "bf task: descriptor pointers must be 16-byte aligned (8 for 4-bit)");
    RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(t.bestMatchA) | reinterpret_cast<uintptr_t>(t.bestMatchB)) % 8 != 0,
                   RSFM_ERR_INVALID, "bf task: result pointers must be 8-byte aligned");
    return RSFM_OK;
}

static DevTask to_dev(const rsfm_bf_task& t)
{
    DevTask d{};
    d.nA = t.nbDescA;
    d.nB = t.nbDescB;
    d.A = static_cast<const uint8_t*>(t.descA);
    d.B = static_cast<const uint8_t*>(t.descB);
    d.bestA = reinterpret_cast<uint64_t*>(t.bestMatchA);
    d.bestB = reinterpret_cast<uint64_t*>(t.bestMatchB);
    return d;
}

// Top-1 entry points: the tile needs the train sets' staging form, which the
// reference interface has no workspace for, so the call takes it from the
// stream-ordered allocator (hipMallocAsync / hipFreeAsync on the same stream:
// a pooled block, no device-wide synchronisation).
template <int kBits>
static int launch_top1(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream)
{
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "bf match: tasks is null");
    for (size_t i = 0; i < nbTasks; i++) RSFM_CHECK(validate_task(tasks[i], kBits));
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        TaskBatch batch{};
        for (size_t i = 0; i < n; i++) batch.t[i] = to_dev(tasks[b + i]);
        const Plan p = make_plan(batch.t, n, kMaxSegColsTop2);
        size_t bytes = 0;
        for (size_t i = 0; i < n; i++) {
            batch.t[i].nSeg = p.nSeg;
            batch.t[i].segCols = p.segCols;
            bytes += prep_bytes(batch.t[i].nA, batch.t[i].nB);
        }
        const uint32_t prepThreads = std::max(div_up(p.maxB, (uint32_t)kBN) * kBN * 8, std::max(p.maxA, p.maxB));
        if (prepThreads == 0) continue;
        void* ws = nullptr;
        if (bytes > 0) RSFM_CHECK(check_hip(hipMallocAsync(&ws, bytes, stream), "bf match: staging workspace"));
        uint8_t* w = static_cast<uint8_t*>(ws);
        for (size_t i = 0; i < n; i++) {
            place_prep(batch.t[i], w);
            w += prep_bytes(batch.t[i].nA, batch.t[i].nB);
        }
        prep_train_kernel<kBits><<<dim3(div_up(prepThreads, 256), 1, (uint32_t)n), 256, 0, stream>>>(batch, kInitTop1);
        int rc = check_launch("bf prep");
        if (rc == RSFM_OK && p.gridX > 0 && p.maxB > 0) {
            match_top1_kernel<kBits><<<dim3(p.gridX, p.nSeg, (uint32_t)n), 512, 0, stream>>>(batch);
            rc = check_launch("bf match");
        }
        if (ws != nullptr) {
            const int rf = check_hip(hipFreeAsync(ws, stream), "bf match: staging workspace release");
            if (rc == RSFM_OK) rc = rf;
        }
        RSFM_CHECK(rc);
    }
    return RSFM_OK;
}

static size_t top2_workspace(const rsfm_bf_top2_task* tasks, size_t nbTasks)
{
    size_t bytes = 0;
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        DevTask t[kMaxTasks];
        for (size_t i = 0; i < n; i++) t[i] = to_dev(tasks[b + i].base);
        const Plan p = make_plan(t, n, kMaxSegColsTop2);
        for (size_t i = 0; i < n; i++) {
            if (p.nSeg > 1) bytes += align_up((size_t)p.nSeg * t[i].nA * 16, 256);
            bytes += prep_bytes(t[i].nA, t[i].nB);
        }
    }
    return bytes;
}

int launch_match_tiles_top2(const DevTask* tasks, const TileDesc* tiles, const uint32_t* xcdOff, uint32_t grid,
                            hipStream_t stream)
{
    RSFM_RETURN_IF(grid == 0 || grid % kNumXcd != 0, RSFM_ERR_INVALID, "match tiles: grid must be a multiple of 8");
    match_tiles_kernel<<<grid, 512, 0, stream>>>(tasks, tiles, xcdOff);
    return check_launch("bf match tiles");
}

int launch_match_tile_grid_top2(const DevTask* tasks, const TileDesc* tiles, const uint32_t* xcdOff, uint32_t maxList,
                                hipStream_t stream)
{
    if (maxList == 0) return RSFM_OK;
    match_tile_grid_kernel<<<maxList * kNumXcd, 512, 0, stream>>>(tasks, tiles, xcdOff);
    return check_launch("bf match tile grid");
}

int launch_merge_top2_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxA, hipStream_t stream)
{
    if (nTasks == 0 || maxA == 0) return RSFM_OK;
    merge_top2_kernel<<<dim3(div_up(maxA, 256), 1, nTasks), 256, 0, stream>>>(TaskTable{tasks});
    return check_launch("bf merge table");
}

}  // namespace bf
}  // namespace rsfm_amd

using namespace rsfm_amd;

extern "C" int rsfm_bf_match_u8(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream)
{
    return bf::launch_top1<8>(tasks, nbTasks, stream);
}

extern "C" int rsfm_bf_match_u4(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream)
{
    return bf::launch_top1<4>(tasks, nbTasks, stream);
}

extern "C" int rsfm_bf_match_top2_workspace_size(const rsfm_bf_top2_task* tasks, size_t nbTasks, size_t* bytes)
{
    RSFM_RETURN_IF(bytes == nullptr, RSFM_ERR_INVALID, "bytes is null");
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "tasks is null");
    *bytes = bf::top2_workspace(tasks, nbTasks);
    return RSFM_OK;
}

extern "C" int rsfm_bf_match_top2(const rsfm_bf_top2_task* tasks, size_t nbTasks, void* workspace,
                                  size_t workspaceBytes, hipStream_t stream)
{
    using namespace bf;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "bf top2: tasks is null");
    for (size_t i = 0; i < nbTasks; i++) {
        RSFM_CHECK(validate_task(tasks[i].base, 8));
        RSFM_RETURN_IF(tasks[i].base.nbDescA > 0 && tasks[i].secondMatchA == nullptr, RSFM_ERR_INVALID,
                       "bf top2: secondMatchA must be set when nbDescA > 0");
        RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(tasks[i].secondMatchA) % 8 != 0, RSFM_ERR_INVALID,
                       "bf top2: secondMatchA must be 8-byte aligned");
    }
    const size_t need = top2_workspace(tasks, nbTasks);
    RSFM_RETURN_IF(workspaceBytes < need, RSFM_ERR_INVALID,
                   "bf top2: workspace too small (" + std::to_string(workspaceBytes) + " < " + std::to_string(need) + ")");
    RSFM_RETURN_IF(need > 0 && (workspace == nullptr || reinterpret_cast<uintptr_t>(workspace) % 256 != 0),
                   RSFM_ERR_INVALID, "bf top2: workspace must be non-null and 256-byte aligned");
    uint8_t* ws = static_cast<uint8_t*>(workspace);
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        TaskBatch batch{};
        for (size_t i = 0; i < n; i++) {
            batch.t[i] = to_dev(tasks[b + i].base);
            batch.t[i].secondA = reinterpret_cast<uint64_t*>(tasks[b + i].secondMatchA);
        }
        const Plan p = make_plan(batch.t, n, kMaxSegColsTop2);
        for (size_t i = 0; i < n; i++) {
            batch.t[i].nSeg = p.nSeg;
            batch.t[i].segCols = p.segCols;
        }
        for (size_t i = 0; i < n; i++) {
            if (p.nSeg > 1) {
                batch.t[i].part = reinterpret_cast<uint64_t*>(ws);
                ws += align_up((size_t)p.nSeg * batch.t[i].nA * 16, 256);
            }
            place_prep(batch.t[i], ws);
            ws += prep_bytes(batch.t[i].nA, batch.t[i].nB);
        }
        // one launch: results to "no match" where no tile writes them, and
        // the staging form of every task's train descriptors
        const uint32_t prepThreads = std::max(div_up(p.maxB, (uint32_t)kBN) * kBN * 8, std::max(p.maxA, p.maxB));
        if (prepThreads == 0) continue;
        prep_train_kernel<8><<<dim3(div_up(prepThreads, 256), 1, (uint32_t)n), 256, 0, stream>>>(batch, kInitTop2);
        RSFM_CHECK(check_launch("bf top2 prep"));
        if (p.gridX == 0 || p.maxB == 0) continue;
        match_top2_kernel<<<dim3(p.gridX, p.nSeg, (uint32_t)n), 512, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("bf top2 match"));
        if (p.nSeg > 1) {
            merge_top2_kernel<<<dim3(div_up(p.maxA, 256), 1, (uint32_t)n), 256, 0, stream>>>(batch);
            RSFM_CHECK(check_launch("bf top2 merge"));
        }
    }
    return RSFM_OK;
}

// ---------------------------------------------------------------------------
// rsfm_bf_match_top2_noalloc: the top-2 matcher with its workspace (train
// staging form and segment partials) in a fixed device array instead of a
// caller buffer, for callers that cannot allocate. The array is this file's
// `scratch`; its device address is looked up once per device, with that device
// current, and only that address is handed to kernels. Every call on one GPU
// shares the array (calls that overlap on one GPU race); nothing in it needs
// clearing (the preparation kernel writes the whole staging form and every
// partial that the merge reads is written by a tile first).
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratchLock;
void* g_scratchPtr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

static int noalloc_scratch(void** out)
{
    int dev = 0;
    RSFM_CHECK(check_hip(hipGetDevice(&dev), "top2 noalloc: hipGetDevice"));
    RSFM_RETURN_IF(dev < 0 || dev >= kMaxDevices, RSFM_ERR_UNSUPPORTED, "top2 noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratchLock);
    if (g_scratchPtr[dev] == nullptr) {
        void* p = nullptr;
        RSFM_CHECK(check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "top2 noalloc: hipGetSymbolAddress"));
        RSFM_RETURN_IF(p == nullptr || reinterpret_cast<uintptr_t>(p) % 256 != 0, RSFM_ERR_HIP,
                       "top2 noalloc: scratch address unusable");
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return RSFM_OK;
}

extern "C" int rsfm_bf_match_top2_noalloc(const rsfm_bf_top2_task* tasks, size_t nbTasks, hipStream_t stream)
{
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "bf top2 noalloc: tasks is null");
    const size_t need = bf::top2_workspace(tasks, nbTasks);
    RSFM_RETURN_IF(need > kNoAllocScratchBytes, RSFM_ERR_INVALID,
                   "bf top2 noalloc: these tasks need " + std::to_string(need) +
                       " bytes of workspace, the fixed scratch holds " + std::to_string(kNoAllocScratchBytes) +
                       " (64 MiB); use rsfm_bf_match_top2 with a caller workspace");
    void* ws = nullptr;
    if (need > 0) RSFM_CHECK(noalloc_scratch(&ws));
    return rsfm_bf_match_top2(tasks, nbTasks, ws, need, stream);
}
