// bf_match.hip - brute-force SIFT matching on gfx950 int8 matrix cores.
//
// Contract (reference: kernelBruteForceMatchIMMA, bruteForceMatchIMMA.cu:551-586,
// packed keys bruteForceMatch.cuh:108-261): for every query i the lexicographic
// minimum of (L2^2(i,j), j) over all train j, and for every train j the minimum
// of (L2^2(i,j), i) over all queries i. Distances are exact integers.
//
// Design (own, for CDNA4):
//  * u8 descriptors are mapped to signed bytes u = 127 - a (rows) and
//    w = b - 128 (columns) so one v_mfma_i32_32x32x32_i8 chain gives
//    dot = u.w and  L2^2 = NA' + NB' + 2 dot  with NA' = |a-128|^2 and
//    NB' = |b-127|^2 - 128 (both in [-128, 2^21]).
//  * A workgroup keeps 256 query rows resident in registers (kRB 32-row MFMA
//    blocks per wave, K = 128 = 4 MFMAs each) and streams the train
//    descriptors of its column segment through LDS in 64-column chunks
//    (double buffered, XOR-swizzled 16-B slots).
//  * Epilogue per accumulator: one v_lshl_add makes the row key
//    ((NB'_j + 2dot) << 8 | column-block tag), one makes the column key
//    ((NA'_i + 2dot) << 8 | local row); v_min3 folds top-1, v_med3 + v_min
//    fold top-2. Row state lives in registers for the whole segment and is
//    reduced at the end through an LDS transpose (one thread per row).
//    Column partials are reduced across the waves through LDS and merged
//    globally with one 64-bit atomicMin per column per workgroup.
//  * Segments split the train range when one launch has too few row blocks
//    to fill 256 CUs; row results then merge by 64-bit atomicMin (top-1) or
//    through a small workspace and a merge kernel (top-2).
#include "internal.hpp"
#include <algorithm>
#include <cstring>
#include <vector>

namespace rsfm_amd {
namespace bf {

constexpr int32_t kNone = INT32_MAX;

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
// The same as asm. The row and the column key shift the same accumulator; with
// both in plain code the compiler shares one shift and adds twice (three
// instructions for two keys), so the column key uses this form.
__device__ inline int32_t shl9_add_asm(int32_t a, int32_t b)
{
    int32_t r;
    asm("v_lshl_add_u32 %0, %1, 9, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
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

// sCol row stride in words: the flush reads 8 consecutive columns from each of
// the 8 waves; a stride of 72 puts those 64 words in distinct banks
constexpr int kColStride = 72;

// One workgroup = 8 waves x 32 query rows. The column segment streams through
// LDS in 64-column chunks. The loop is software pipelined: in iteration c the
// wave issues the 8 MFMAs of chunk c into one accumulator pair while the VALU
// folds the keys of chunk c-1 from the other pair, so matrix and vector work of
// one wave overlap; the body has no branches, so the scheduler can interleave
// them. Column partials of chunk c-1 are written to LDS in iteration c and
// merged to global memory by all waves (8 columns each) in iteration c+1.
template <int kBits, bool kTop2>
__device__ __forceinline__ void match_tile(const DevTask& T, uint32_t rowBase, uint32_t colBeg, uint32_t seg)
{
    constexpr int kWaves = 8;
    constexpr int kDescBytes = kBits * 16;
    constexpr int kChunkBytes = kBN * 128;

    const uint32_t colEnd = min(colBeg + T.segCols, T.nB);
    const uint32_t nChunks = (colEnd - colBeg + kBN - 1) / kBN;

    // ring of staged chunks; the row-tail transpose reuses its memory
    constexpr int kRing = 4, kColSlots = 4, kPSlots = 8;
    constexpr int kTailStride = 33;
    constexpr int kUnionBytes = (kRing * kChunkBytes > kBM * kTailStride * 4) ? kRing * kChunkBytes : kBM * kTailStride * 4;
    __shared__ __attribute__((aligned(16))) uint8_t sU[kUnionBytes];
    __shared__ int32_t sP[kPSlots][kBN];  // per column: NB' << 8 | column block (kNone past the end)
    __shared__ int32_t sCol[kColSlots][kWaves][kColStride];
    __shared__ int32_t sNA[kBM];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, hi = lane >> 5;

    // ---- query rows: operands u = 127 - a, offsets NA' = |a - 128|^2
    v4i aop[4];
    int32_t Q[16];
    {
        const int rloc = wave * 32 + l32;
        const uint32_t row = rowBase + rloc;
        const bool valid = row < T.nA;
        int32_t nap = 0;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            v4i x = {0, 0, 0, 0};
            if (valid) x = load16<kBits>(T.A + (size_t)row * kDescBytes, kc * 32 + hi * 16);
            v4i u;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int xa = x[q] ^ (int)0x80808080u;
                nap = __builtin_amdgcn_sdot4(xa, xa, nap, false);
                u[q] = valid ? (x[q] ^ 0x7F7F7F7F) : 0;
            }
            aop[kc] = u;
        }
        const int32_t na = nap + __shfl_xor(nap, 32);
        if (hi == 0) sNA[rloc] = na;
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int r = 8 * (v >> 2) + 4 * hi + (v & 3);
            const int32_t nav = __shfl(na, r);
            const int rl = wave * 32 + r;
            Q[v] = rowBase + rl < T.nA ? (nav * 256 + rl) : kNone;
        }
    }

    int32_t m1[16], m2[16];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        m1[v] = kNone;
        m2[v] = kNone;
    }

    // ---- staging: thread -> 16 bytes (slot sp) of chunk column sc. Chunks live
    // in a ring of kRing LDS slots; chunk c is staged while chunk c-2 is being
    // processed and the workgroup synchronises after every odd chunk. Between
    // the staging of a slot and its next read, and between its read and its
    // next staging, there is then always one barrier.
    const int sc = tid >> 3, sp = tid & 7;
    auto stage_load = [&](uint32_t c) -> v4i {
        // past the segment end: any valid column (the value is masked at the store)
        const uint32_t gcol = min(colBeg + c * kBN + sc, colEnd - 1);
        return load16<kBits>(T.B + (size_t)gcol * kDescBytes, sp * 16);
    };
    auto stage_store = [&](uint32_t c, const v4i& x) {
        const bool valid = colBeg + c * kBN + sc < colEnd;
        int32_t acc = 0;
        v4i w;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            w[q] = valid ? (x[q] ^ (int)0x80808080u) : 0;
            acc = __builtin_amdgcn_sdot4(w[q], w[q], acc, false);
            acc = __builtin_amdgcn_sdot4(w[q], 0x02020202, acc, false);
        }
        *reinterpret_cast<v4i*>(sU + (c % kRing) * kChunkBytes + sc * 128 + ((sp ^ swz(sc)) << 4)) = w;
        // sum over the 8 threads of the column: (w^2 + 2w) over 128 values = |b - 127|^2 - 128 = NB'
        acc += dpp_xor1(acc);
        acc += dpp_xor2(acc);
        acc += dpp_mirror8(acc);
        // row-key offset NB' << 8 | column block; columns past the end never win
        sP[c % kPSlots][sc] = valid ? (acc * 256 + (int32_t)(c * 2 + (sc >> 5))) : kNone;  // 8 threads, same value
    };
    // merged column minima of chunk k -> one 64-bit atomicMin per column (8 columns per wave)
    auto flush = [&](uint32_t k) {
        const int col = wave * 8 + (lane >> 3);
        int32_t x = sCol[k % kColSlots][lane & 7][col];
        x = imin(x, dpp_xor1(x));
        x = imin(x, dpp_xor2(x));
        x = imin(x, dpp_mirror8(x));
        const uint32_t gcol = colBeg + k * kBN + col;
        if ((lane & 7) == 0 && gcol < colEnd && x != kNone) {
            const uint32_t dist = (uint32_t)((x >> 8) + (sP[k % kPSlots][col] >> 8));
            const uint32_t row = rowBase + (uint32_t)(x & 0xFF);
            atomicMin(reinterpret_cast<unsigned long long*>(T.bestB + gcol), (unsigned long long)make_key(dist, row));
        }
    };
    // keys of row v of one chunk: the row keys (NB'_j + 2 dot) << 8 | column
    // block fold into the row state, the column keys (NA'_i + 2 dot) << 8 |
    // local row into the running column minima cm0, cm1 of the lane
    auto fold_row = [&](int v, const v16i (&acc)[2], const int32_t (&P)[2], int32_t& cm0, int32_t& cm1) {
        const int32_t a0 = acc[0][v], a1 = acc[1][v];
        const int32_t r0 = shl9_add(a0, P[0]);
        const int32_t r1 = shl9_add(a1, P[1]);
        if constexpr (kTop2)
            top2_fold2(m1[v], m2[v], r0, r1);
        else
            m1[v] = imin(m1[v], imin(r0, r1));
        cm0 = imin(cm0, shl9_add_asm(a0, Q[v]));
        cm1 = imin(cm1, shl9_add_asm(a1, Q[v]));
    };
    // lanes l and l + 32 hold the same column: exchange the upper half of
    // block 0 with the lower half of block 1, then lane l holds column l
    auto put_cols = [&](int32_t cm0, int32_t cm1, uint32_t k) {
        const auto sw = __builtin_amdgcn_permlane32_swap((uint32_t)cm0, (uint32_t)cm1, false, false);
        sCol[k % kColSlots][wave][lane] = imin((int32_t)sw[0], (int32_t)sw[1]);
    };
    auto fold_keys = [&](const v16i (&acc)[2], const int32_t (&P)[2], uint32_t k) {
        int32_t cm0 = kNone, cm1 = kNone;
#pragma unroll
        for (int v = 0; v < 16; v++) fold_row(v, acc, P, cm0, cm1);
        put_cols(cm0, cm1, k);
    };
    const bool doCols = T.bestB != nullptr;

    // chunk c: MFMAs of chunk c -> accN (row offsets Pn); keys of chunk c-1 from
    // accP (row offsets Pp); stages chunk c+2 from xs and loads chunk c+4 into xs
    auto body = [&](uint32_t c, v16i (&accN)[2], int32_t (&Pn)[2], const v16i (&accP)[2], const int32_t (&Pp)[2],
                    v4i& xs, bool sync) {
        stage_store(c + 2, xs);
        const uint8_t* sb = sU + (c % kRing) * kChunkBytes;
        v4i bop[2][4];
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int kc = 0; kc < 4; kc++) {
                const int cl = cb * 32 + l32;
                const int s = 2 * kc + hi;
                bop[cb][kc] = *reinterpret_cast<const v4i*>(sb + cl * 128 + ((s ^ swz(cl)) << 4));
            }
        xs = stage_load(c + 4);
        // the keys of rows 2i, 2i+1 of the previous chunk, then MFMA i: the
        // scheduling fences keep that order, so each wave overlaps its own
        // vector and matrix work, the first MFMA finds its LDS operand loaded
        // and dependent MFMAs are ~18 VALU apart
        int32_t cm0 = kNone, cm1 = kNone;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int cb = i >> 2, kc = i & 3;
            fold_row(2 * i, accP, Pp, cm0, cm1);
            fold_row(2 * i + 1, accP, Pp, cm0, cm1);
            __builtin_amdgcn_sched_barrier(0);
            const v16i zero = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            accN[cb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aop[kc], bop[cb][kc], kc == 0 ? zero : accN[cb], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        put_cols(cm0, cm1, c - 1);
#pragma unroll
        for (int cb = 0; cb < 2; cb++) Pn[cb] = sP[c % kPSlots][cb * 32 + l32];
        // keep this chunk's row folding here (the compiler would otherwise
        // sink it past the barrier into the next chunk)
#pragma unroll
        for (int v = 0; v < 16; v++) asm volatile("" : "+v"(m1[v]), "+v"(m2[v]));
        if (sync) __syncthreads();
    };
    // column partials of chunks k-3 and k-2 are complete after the barrier that
    // precedes the even chunk k
    auto flush_before = [&](uint32_t k) {
        if (!doCols) return;
        if (k >= 3) flush(k - 3);
        if (k >= 2) flush(k - 2);
    };

    v16i accA[2], accB[2];
    int32_t PA[2], PB[2];
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        accB[cb] = v16i{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        PB[cb] = kNone;  // the keys of "chunk -1" never win
    }
    v4i xs0 = stage_load(0), xs1 = stage_load(1);
    stage_store(0, xs0);
    stage_store(1, xs1);
    xs0 = stage_load(2);
    xs1 = stage_load(3);
    __syncthreads();

    // two chunks per trip: the accumulator pairs swap roles without copies and
    // the barrier follows the odd chunk
    uint32_t c = 0;
    for (; c + 2 <= nChunks; c += 2) {
        flush_before(c);
        body(c, accA, PA, accB, PB, xs0, false);
        body(c + 1, accB, PB, accA, PA, xs1, true);
    }
    uint32_t lastEven;  // the last even chunk
    if (c < nChunks) {
        flush_before(c);
        body(c, accA, PA, accB, PB, xs0, false);
        fold_keys(accA, PA, nChunks - 1);
        lastEven = c;
    } else {
        fold_keys(accB, PB, nChunks - 1);
        lastEven = c - 2;
    }
    __syncthreads();
    // chunks from lastEven - 1 on are not flushed yet
    if (doCols)
        for (uint32_t k = lastEven >= 1 ? lastEven - 1 : 0; k < nChunks; k++) flush(k);

    // ---- row results: transpose the per-lane (row, column-class) bests through LDS
    int32_t* L = reinterpret_cast<int32_t*>(sU);
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int rl = wave * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
        L[rl * kTailStride + l32] = m1[v];
    }
    __syncthreads();
    // thread tid owns local row tid
    int32_t b1 = kNone, b2 = kNone;
    int c1 = 0, c2 = 0;
    if (tid < kBM) {
        for (int cc = 0; cc < 32; cc++) {
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
        for (int v = 0; v < 16; v++) {
            const int rl = wave * 32 + 8 * (v >> 2) + 4 * hi + (v & 3);
            L[rl * kTailStride + l32] = m2[v];
        }
        __syncthreads();
        if (tid < kBM) {
            // second best = min(second best over classes, second best inside the winning class)
            const int32_t x = L[tid * kTailStride + c1];
            if (x < b2 || (x == b2 && c1 < c2)) {
                b2 = x;
                c2 = c1;
            }
        }
    }
    if (tid < kBM) {
        const uint32_t row = rowBase + tid;
        if (row < T.nA) {
            const int32_t na = sNA[tid];
            auto to_key = [&](int32_t m, int cls) -> uint64_t {
                if (m == kNone) return kNoMatch;
                return make_key((uint32_t)((m >> 8) + na), colBeg + (uint32_t)(m & 0xFF) * 32u + (uint32_t)cls);
            };
            const uint64_t k1 = to_key(b1, c1);
            const bool direct = T.nSeg == 1;
            if constexpr (kTop2) {
                const uint64_t k2 = to_key(b2, c2);
                if (direct) {
                    T.bestA[row] = k1;
                    T.secondA[row] = k2;
                } else {
                    uint64_t* p = T.part + ((size_t)seg * T.nA + row) * 2;
                    p[0] = k1;
                    p[1] = k2;
                }
            } else {
                if (direct)
                    T.bestA[row] = k1;
                else if (k1 != kNoMatch)
                    atomicMin(reinterpret_cast<unsigned long long*>(T.bestA + row), (unsigned long long)k1);
            }
        }
    }
}

// grid launch: x = row block, y = segment, z = task (XCD-local order)
template <int kBits, bool kTop2>
__global__ __launch_bounds__(512) void match_kernel(const TaskBatch batch)
{
    const Wg3 wg = xcd_local_block();
    const DevTask& T = batch.t[wg.z];
    const uint32_t rowBase = wg.x * kBM;
    const uint32_t colBeg = wg.y * T.segCols;
    if (rowBase >= T.nA || colBeg >= T.nB) return;
    match_tile<kBits, kTop2>(T, rowBase, colBeg, wg.y);
}

// Persistent launch over a tile table: the tiles are split into kNumXcd lists
// (xcdOff[x] .. xcdOff[x+1]), one per XCD; workgroup b serves list b % kNumXcd
// (the XCD the hardware dispatches it to), striding by the workgroups of that
// XCD, so all tiles of one task share one L2. gridDim.x is a multiple of
// kNumXcd.
template <int kBits, bool kTop2>
__global__ __launch_bounds__(512) void match_tiles_kernel(const DevTask* __restrict__ tasks,
                                                          const TileDesc* __restrict__ tiles,
                                                          const uint32_t* __restrict__ xcdOff)
{
    const uint32_t xcd = blockIdx.x % kNumXcd, per = gridDim.x / kNumXcd;
    const uint32_t end = xcdOff[xcd + 1];
    for (uint32_t i = xcdOff[xcd] + blockIdx.x / kNumXcd; i < end; i += per) {
        const TileDesc d = tiles[i];
        const DevTask T = tasks[d.task];
        __syncthreads();  // the previous tile's last LDS reads
        match_tile<kBits, kTop2>(T, d.rowBase, d.colBeg, d.seg);
    }
}

// The same tile table, one workgroup per tile: workgroup b takes element b / 8
// of XCD list b % 8 (the XCD the hardware dispatches it to). The body is
// match_tile alone, without the persistent loop's extra live registers.
template <int kBits, bool kTop2>
__global__ __launch_bounds__(512) void match_tile_grid_kernel(const DevTask* __restrict__ tasks,
                                                              const TileDesc* __restrict__ tiles,
                                                              const uint32_t* __restrict__ xcdOff)
{
    const uint32_t xcd = blockIdx.x % kNumXcd;
    const uint32_t i = xcdOff[xcd] + blockIdx.x / kNumXcd;
    if (i >= xcdOff[xcd + 1]) return;
    const TileDesc d = tiles[i];
    const DevTask T = tasks[d.task];
    match_tile<kBits, kTop2>(T, d.rowBase, d.colBeg, d.seg);
}

template <class Src>
__global__ void init_kernel(const Src batch, bool initA)
{
    const DevTask& T = batch[blockIdx.z];
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (initA && x < T.nA) T.bestA[x] = kNoMatch;
    if (T.bestB != nullptr && x < T.nB) T.bestB[x] = kNoMatch;
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

static Plan make_plan(const DevTask* t, size_t n)
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
    nSeg = std::max(nSeg, div_up(p.maxB, kMaxSegCols));
    uint32_t seg = div_up(p.maxB, nSeg);
    seg = div_up(seg, kBN) * kBN;
    p.segCols = std::min(seg, kMaxSegCols);
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

template <int kBits>
static int launch_top1(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream)
{
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "bf match: tasks is null");
    for (size_t i = 0; i < nbTasks; i++) RSFM_CHECK(validate_task(tasks[i], kBits));
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        TaskBatch batch{};
        for (size_t i = 0; i < n; i++) batch.t[i] = to_dev(tasks[b + i]);
        const Plan p = make_plan(batch.t, n);
        for (size_t i = 0; i < n; i++) {
            batch.t[i].nSeg = p.nSeg;
            batch.t[i].segCols = p.segCols;
        }
        const uint32_t initX = div_up(std::max(p.maxA, p.maxB), 256);
        if (initX == 0) continue;
        init_kernel<<<dim3(initX, 1, (uint32_t)n), 256, 0, stream>>>(batch, true);
        RSFM_CHECK(check_launch("bf init"));
        if (p.gridX == 0 || p.maxB == 0) continue;
        match_kernel<kBits, false><<<dim3(p.gridX, p.nSeg, (uint32_t)n), 512, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("bf match"));
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
        const Plan p = make_plan(t, n);
        if (p.nSeg <= 1) continue;
        for (size_t i = 0; i < n; i++) bytes += align_up((size_t)p.nSeg * t[i].nA * 16, 256);
    }
    return bytes;
}

int launch_match_tiles_top2(const DevTask* tasks, const TileDesc* tiles, const uint32_t* xcdOff, uint32_t grid,
                            hipStream_t stream)
{
    RSFM_RETURN_IF(grid == 0 || grid % kNumXcd != 0, RSFM_ERR_INVALID, "match tiles: grid must be a multiple of 8");
    match_tiles_kernel<8, true><<<grid, 512, 0, stream>>>(tasks, tiles, xcdOff);
    return check_launch("bf match tiles");
}

int launch_match_tile_grid_top2(const DevTask* tasks, const TileDesc* tiles, const uint32_t* xcdOff, uint32_t maxList,
                                hipStream_t stream)
{
    if (maxList == 0) return RSFM_OK;
    match_tile_grid_kernel<8, true><<<maxList * kNumXcd, 512, 0, stream>>>(tasks, tiles, xcdOff);
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
        const Plan p = make_plan(batch.t, n);
        for (size_t i = 0; i < n; i++) {
            batch.t[i].nSeg = p.nSeg;
            batch.t[i].segCols = p.segCols;
        }
        if (p.nSeg > 1) {
            for (size_t i = 0; i < n; i++) {
                batch.t[i].part = reinterpret_cast<uint64_t*>(ws);
                ws += align_up((size_t)p.nSeg * batch.t[i].nA * 16, 256);
            }
        }
        const uint32_t initX = div_up(std::max(p.maxA, p.maxB), 256);
        if (initX == 0) continue;
        // with no train descriptors the query side must still read "no match"
        init_kernel<<<dim3(initX, 1, (uint32_t)n), 256, 0, stream>>>(batch, true);
        RSFM_CHECK(check_launch("bf top2 init"));
        if (p.maxA > 0) {
            TaskBatch b2 = batch;
            for (size_t i = 0; i < n; i++) {
                b2.t[i].bestA = batch.t[i].secondA;
                b2.t[i].bestB = nullptr;
            }
            init_kernel<<<dim3(div_up(p.maxA, 256), 1, (uint32_t)n), 256, 0, stream>>>(b2, true);
            RSFM_CHECK(check_launch("bf top2 init2"));
        }
        if (p.gridX == 0 || p.maxB == 0) continue;
        match_kernel<8, true><<<dim3(p.gridX, p.nSeg, (uint32_t)n), 512, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("bf top2 match"));
        if (p.nSeg > 1) {
            merge_top2_kernel<<<dim3(div_up(p.maxA, 256), 1, (uint32_t)n), 256, 0, stream>>>(batch);
            RSFM_CHECK(check_launch("bf top2 merge"));
        }
    }
    return RSFM_OK;
}
