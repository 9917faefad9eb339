// mlp.hip - the learned RANSAC proposal network on gfx950 bf16 matrix cores.
//
// Network (no reference counterpart; the reference samples uniformly,
// ransac.hpp:23-82): 8 inputs -> 64 -> 64 -> 64 (ReLU) -> 1 (sigmoid), mean
// binary cross-entropy. Parameters are an FP32 master vector
//   W1[64x8] b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[64] b4
// (row-major W[out][in]); the kernels round weights and inputs to bf16 (round
// to nearest even) and accumulate in FP32.
//
// Layout (own design): the network is evaluated transposed, H^T = W X^T, so
// a v_mfma_f32_32x32x16_bf16 output block holds one match per lane column and
// 16 hidden units per lane. The hidden units of each 16-wide K slice are
// permuted (sigma below) so that the ReLU'd output of one layer, packed to
// bf16, is directly the B operand of the next layer: activations never leave
// registers in the forward pass. The weight operands are re-laid out once per
// workgroup into LDS in exactly the order the lanes read them.
//
// Backward: deltas flow through the transposed weights with the same trick;
// the weight gradients dW = G H^T contract over the batch, so deltas and
// activations of 128 matches are transposed through LDS ([unit][match]) and
// fed to bf16 MFMAs whose K dimension is the match index. Biases ride along
// as a column of ones. Each workgroup writes one partial gradient vector;
// a second kernel sums the partials in a fixed order (deterministic).
#include "common.hpp"
#include <algorithm>
#include <string>

namespace rsfm_amd {
namespace mlp {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kH = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + kH * 8;
constexpr int OFF_W2 = OFF_B1 + kH;
constexpr int OFF_B2 = OFF_W2 + kH * kH;
constexpr int OFF_W3 = OFF_B2 + kH;
constexpr int OFF_B3 = OFF_W3 + kH * kH;
constexpr int OFF_W4 = OFF_B3 + kH;
constexpr int OFF_B4 = OFF_W4 + kH;
constexpr int kNParam = OFF_B4 + 1;

constexpr int kWaves = 4;               // 256 threads
constexpr int kStep = kWaves * 32;      // matches per workgroup step
constexpr int kRowStride = kStep + 8;   // bf16 elements per LDS row of a transposed array

// K slot s (0..15) of a 16-wide slice <-> hidden unit 16 t + sigma(s) of the
// previous layer, the unit that sits in lane half s / 8, register (s % 8) of
// the previous layer's output block (rows 8 (v / 4) + 4 hi + v % 4)
__device__ inline int sigma(int s) { return 8 * ((s & 7) >> 2) + 4 * (s >> 3) + (s & 3); }

__device__ inline __bf16 to_bf(float f) { return (__bf16)f; }
__device__ inline float from_bf(__bf16 h) { return (float)h; }

// operand tables in LDS (one v8bf per lane per (row block, K slice))
struct Ops {
    v8bf w1[2][64];
    v8bf w2[2][4][64];
    v8bf w3[2][4][64];
    v8bf w3t[2][4][64];
    v8bf w2t[2][4][64];
    float bias[3][2][2][16];  // [layer][block][hi][v] in output-block order
    float w4[2][32];          // [hi][16 b + v], bf16-rounded
    float b4;
};

__device__ inline v16f zero16() { return v16f{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}; }

// build the operand tables from the FP32 master parameters
__device__ void build_ops(Ops& S, const float* __restrict__ P)
{
    const int tid = threadIdx.x;
    for (int e = tid; e < 2 * 64; e += blockDim.x) {
        const int rb = e >> 6, lane = e & 63, m = lane & 31, hi = lane >> 5;
        v8bf r;
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = hi ? to_bf(0.f) : to_bf(P[OFF_W1 + (32 * rb + m) * 8 + j]);
        S.w1[rb][lane] = r;
    }
    for (int e = tid; e < 2 * 4 * 64; e += blockDim.x) {
        const int rb = e >> 8, t = (e >> 6) & 3, lane = e & 63, m = lane & 31, hi = lane >> 5;
        v8bf a2, a3, a3t, a2t;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int u = 16 * t + sigma(8 * hi + j);  // input unit of the forward layer
            const int r = 32 * rb + m;                // output row of this operand
            a2[j] = to_bf(P[OFF_W2 + r * kH + u]);
            a3[j] = to_bf(P[OFF_W3 + r * kH + u]);
            a3t[j] = to_bf(P[OFF_W3 + u * kH + r]);
            a2t[j] = to_bf(P[OFF_W2 + u * kH + r]);
        }
        S.w2[rb][t][lane] = a2;
        S.w3[rb][t][lane] = a3;
        S.w3t[rb][t][lane] = a3t;
        S.w2t[rb][t][lane] = a2t;
    }
    for (int e = tid; e < 3 * 2 * 2 * 16; e += blockDim.x) {
        const int L = e / 64, b = (e >> 5) & 1, hi = (e >> 4) & 1, v = e & 15;
        const int o = 32 * b + 8 * (v >> 2) + 4 * hi + (v & 3);
        const int off = L == 0 ? OFF_B1 : (L == 1 ? OFF_B2 : OFF_B3);
        S.bias[L][b][hi][v] = P[off + o];
    }
    for (int e = tid; e < 64; e += blockDim.x) {
        const int hi = e >> 5, bv = e & 31, b = bv >> 4, v = bv & 15;
        const int o = 32 * b + 8 * (v >> 2) + 4 * hi + (v & 3);
        S.w4[hi][bv] = from_bf(to_bf(P[OFF_W4 + o]));
    }
    if (tid == 0) S.b4 = P[OFF_B4];
}

__device__ inline v16f bias_block(const Ops& S, int L, int b, int hi)
{
    v16f c;
#pragma unroll
    for (int v = 0; v < 16; v++) c[v] = S.bias[L][b][hi][v];
    return c;
}

// ReLU of two output blocks, packed as the four B operands of the next layer
__device__ inline void relu_pack(const v16f (&a)[2], v8bf (&B)[4])
{
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float x = a[t >> 1][8 * (t & 1) + j];
            B[t][j] = to_bf(x > 0.f ? x : 0.f);
        }
}

// one hidden layer: out[rb] = bias + sum_t W[rb][t] * B[t]
__device__ inline void layer64(const v8bf (&W)[2][4][64], const Ops& S, int L, int lane, const v8bf (&B)[4],
                               v16f (&out)[2])
{
    const int hi = lane >> 5;
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        v16f c = bias_block(S, L, rb, hi);
#pragma unroll
        for (int t = 0; t < 4; t++) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[rb][t][lane], B[t], c, 0, 0, 0);
        out[rb] = c;
    }
}

// forward of the 32 matches of one wave tile; returns the logit z (lanes l and
// l + 32 both hold the logit of match l % 32) and leaves the packed
// activations for the backward pass
struct Acts {
    v8bf x;      // layer-1 B operand (features in the low half-wave, zeros above)
    v8bf h1[4];  // layer-2 B operands = bf16(relu(a1))
    v8bf h2[4];
    v8bf h3[4];
};

__device__ inline float forward_tile(const Ops& S, const float* __restrict__ X, uint32_t n, uint32_t base, int lane,
                                     Acts& A)
{
    const int hi = lane >> 5;
    const uint32_t idx = base + (lane & 31);
    v8bf xb;
    if (hi == 0 && idx < n) {
        const float4 p = *reinterpret_cast<const float4*>(X + (size_t)idx * 8);
        const float4 q = *reinterpret_cast<const float4*>(X + (size_t)idx * 8 + 4);
        xb = v8bf{to_bf(p.x), to_bf(p.y), to_bf(p.z), to_bf(p.w), to_bf(q.x), to_bf(q.y), to_bf(q.z), to_bf(q.w)};
    } else {
        const __bf16 z = to_bf(0.f);
        xb = v8bf{z, z, z, z, z, z, z, z};
    }
    A.x = xb;
    v16f a[2];
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
        a[rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.w1[rb][lane], xb, bias_block(S, 0, rb, hi), 0, 0, 0);
    relu_pack(a, A.h1);
    layer64(S.w2, S, 1, lane, A.h1, a);
    relu_pack(a, A.h2);
    layer64(S.w3, S, 2, lane, A.h2, a);
    relu_pack(a, A.h3);
    float z = 0.f;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < 8; j++) z += S.w4[hi][8 * t + j] * from_bf(A.h3[t][j]);
    z += __shfl_xor(z, 32);
    return z + S.b4;
}

__global__ __launch_bounds__(256) void forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                      uint32_t n, float* __restrict__ prob)
{
    __shared__ Ops S;
    build_ops(S, P);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t base = (blockIdx.x * kWaves + wave) * 32u; base < n; base += gridDim.x * kWaves * 32u) {
        Acts A;
        const float z = forward_tile(S, X, n, base, lane, A);
        const uint32_t idx = base + lane;
        if (lane < 32 && idx < n) prob[idx] = 1.f / (1.f + __expf(-z));
    }
}

// transposed copies of one workgroup step ([unit][match], bf16)
struct Trans {
    __bf16 g3[kH][kRowStride];
    __bf16 h2[kH][kRowStride];
    __bf16 g2[kH][kRowStride];
    __bf16 h1[kH][kRowStride];
    __bf16 g1[kH][kRowStride];
    __bf16 h3[kH][kRowStride];
    __bf16 x[8][kRowStride];
    __bf16 dz[kRowStride];
    float loss[kWaves];
};

// store the two C-layout blocks of packed activations (B-operand order) as rows
__device__ inline void store_packed(__bf16 (*dst)[kRowStride], const v8bf (&B)[4], int col, int hi)
{
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < 8; j++) dst[16 * t + sigma(8 * hi + j)][col] = B[t][j];
}

// B operand of a ones column (bias gradient): column 0 of the block is 1
__device__ inline v8bf ones_col(int lane)
{
    const __bf16 o = to_bf((lane & 31) == 0 ? 1.f : 0.f);
    return v8bf{o, o, o, o, o, o, o, o};
}

__device__ inline v8bf row8(const __bf16* row, int k0) { return *reinterpret_cast<const v8bf*>(row + k0); }

// The 17 gradient blocks (32 x 32 each):
//   0..5   layer 3: (ob, ib) = (blk / 3, blk % 3), ib 2 = bias column
//   6..11  layer 2: same
//   12,13  layer 1: ob = blk - 12; B columns 0..7 = inputs, 8 = ones
//   14..16 layer 4: ib = blk - 14, A row 0 = dz
constexpr int kBlocks = 17;
constexpr int kBlocksPerWave = (kBlocks + kWaves - 1) / kWaves;

__device__ inline v16f grad_block_step(const Trans& T, int blk, int lane, v16f acc)
{
    const int m = lane & 31, hi = lane >> 5;
    const __bf16 z0 = to_bf(0.f);
    const v8bf zero8 = {z0, z0, z0, z0, z0, z0, z0, z0};
#pragma unroll 2
    for (int kk = 0; kk < kStep / 16; kk++) {
        const int k0 = 16 * kk + 8 * hi;
        v8bf a, b;
        if (blk < 12) {
            const int L3 = blk < 6;
            const int ob = (blk % 6) / 3, ib = (blk % 6) % 3;
            a = row8(L3 ? T.g3[32 * ob + m] : T.g2[32 * ob + m], k0);
            b = ib == 2 ? ones_col(lane) : row8(L3 ? T.h2[32 * ib + m] : T.h1[32 * ib + m], k0);
        } else if (blk < 14) {
            const int ob = blk - 12;
            a = row8(T.g1[32 * ob + m], k0);
            b = m < 8 ? row8(T.x[m], k0) : (m == 8 ? ones_col(0) : zero8);
        } else {
            const int ib = blk - 14;
            a = m == 0 ? row8(T.dz, k0) : zero8;
            b = ib == 2 ? ones_col(lane) : row8(T.h3[32 * ib + m], k0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// write block blk of the accumulated gradient into the parameter-ordered vector
__device__ inline void grad_block_store(float* __restrict__ out, int blk, int lane, const v16f& acc)
{
    const int col = lane & 31, hi = lane >> 5;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int r = 8 * (v >> 2) + 4 * hi + (v & 3);
        const float g = acc[v];
        if (blk < 12) {
            const int L3 = blk < 6;
            const int ob = (blk % 6) / 3, ib = (blk % 6) % 3;
            const int o = 32 * ob + r;
            if (ib < 2)
                out[(L3 ? OFF_W3 : OFF_W2) + o * kH + 32 * ib + col] = g;
            else if (col == 0)
                out[(L3 ? OFF_B3 : OFF_B2) + o] = g;
        } else if (blk < 14) {
            const int o = 32 * (blk - 12) + r;
            if (col < 8)
                out[OFF_W1 + o * 8 + col] = g;
            else if (col == 8)
                out[OFF_B1 + o] = g;
        } else if (r == 0) {
            const int ib = blk - 14;
            if (ib < 2)
                out[OFF_W4 + 32 * ib + col] = g;
            else if (col == 0)
                out[OFF_B4] = g;
        }
    }
}

// Per-workgroup partial gradient of the mean BCE over n samples (scale invN)
// and partial loss sum: partials[wg][0 .. kNParam) and partials[wg][kNParam].
__global__ __launch_bounds__(256) void grad_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                   const float* __restrict__ Y, uint32_t n, float invN,
                                                   float* __restrict__ partials)
{
    __shared__ Ops S;
    __shared__ Trans T;
    build_ops(S, P);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, hi = lane >> 5;
    const int col = wave * 32 + (lane & 31);  // match column of this lane in a step

    v16f acc[kBlocksPerWave];
#pragma unroll
    for (int j = 0; j < kBlocksPerWave; j++) acc[j] = zero16();
    float lossSum = 0.f;

    for (uint32_t stepBase = blockIdx.x * kStep; stepBase < n; stepBase += gridDim.x * kStep) {
        const uint32_t base = stepBase + wave * 32;
        Acts A;
        const float z = forward_tile(S, X, n, base, lane, A);
        const uint32_t idx = base + (lane & 31);
        const bool valid = idx < n;
        const float y = valid ? Y[idx] : 0.f;
        const float p = 1.f / (1.f + __expf(-z));
        const float dz = valid ? (p - y) * invN : 0.f;
        if (valid && hi == 0) lossSum += fmaxf(z, 0.f) - y * z + log1pf(__expf(-fabsf(z)));

        // deltas: g3 = dz * W4 (masked by h3 > 0), then through W3^T and W2^T
        v8bf g3[4], g2[4], g1[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float h = from_bf(A.h3[t][j]);
                g3[t][j] = to_bf(h > 0.f ? dz * S.w4[hi][8 * t + j] : 0.f);
            }
        v16f c[2];
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            v16f s = zero16();
#pragma unroll
            for (int t = 0; t < 4; t++) s = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.w3t[rb][t][lane], g3[t], s, 0, 0, 0);
            c[rb] = s;
        }
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float h = from_bf(A.h2[t][j]);
                g2[t][j] = to_bf(h > 0.f ? c[t >> 1][8 * (t & 1) + j] : 0.f);
            }
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            v16f s = zero16();
#pragma unroll
            for (int t = 0; t < 4; t++) s = __builtin_amdgcn_mfma_f32_32x32x16_bf16(S.w2t[rb][t][lane], g2[t], s, 0, 0, 0);
            c[rb] = s;
        }
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float h = from_bf(A.h1[t][j]);
                g1[t][j] = to_bf(h > 0.f ? c[t >> 1][8 * (t & 1) + j] : 0.f);
            }

        // transpose this wave's 32 matches into the step arrays
        store_packed(T.g3, g3, col, hi);
        store_packed(T.h2, A.h2, col, hi);
        store_packed(T.g2, g2, col, hi);
        store_packed(T.h1, A.h1, col, hi);
        store_packed(T.g1, g1, col, hi);
        store_packed(T.h3, A.h3, col, hi);
        if (hi == 0) {
#pragma unroll
            for (int j = 0; j < 8; j++) T.x[j][col] = A.x[j];
            T.dz[col] = to_bf(dz);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kBlocksPerWave; j++) {
            const int blk = wave + kWaves * j;
            if (blk < kBlocks) acc[j] = grad_block_step(T, blk, lane, acc[j]);
        }
        __syncthreads();
    }

    float* out = partials + (size_t)blockIdx.x * (kNParam + 1);
#pragma unroll
    for (int j = 0; j < kBlocksPerWave; j++) {
        const int blk = wave + kWaves * j;
        if (blk < kBlocks) grad_block_store(out, blk, lane, acc[j]);
    }
    // loss: wave sum, then workgroup sum
    for (int o = 32; o > 0; o >>= 1) lossSum += __shfl_xor(lossSum, o);
    if (lane == 0) T.loss[wave] = lossSum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int w = 0; w < kWaves; w++) s += T.loss[w];
        out[kNParam] = s;
    }
}

// grad[p] = sum over partials in workgroup order (deterministic); loss = mean
__global__ void reduce_kernel(const float* __restrict__ partials, uint32_t nPart, uint32_t n, float* __restrict__ grad,
                              float* __restrict__ loss)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > (uint32_t)kNParam) return;
    float s = 0.f;
    for (uint32_t w = 0; w < nPart; w++) s += partials[(size_t)w * (kNParam + 1) + p];
    if (p < (uint32_t)kNParam)
        grad[p] = s;
    else if (loss)
        *loss = s / (float)n;
}

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                            float* __restrict__ V, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kNParam) return;
    const float g = G[p];
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// ---- per-match input features of the proposal network
struct FeatTask {
    const uint32_t* matches;  // ValidMatch {idxQuery, idxTrain, float distance} x count
    const uint32_t* count;
    uint32_t maxMatches;
    const uint64_t* second;   // BestMatch per query (distance in the high word)
    const float4* kpQ;        // (x, y, scale, orientation)
    const float4* kpT;
    float invW, invH;
    float* x;                 // [maxMatches][8]
};
constexpr int kFeatMaxTasks = 32;
struct FeatBatch {
    FeatTask t[kFeatMaxTasks];
};

// f0 ratio sqrt(d1/d2) (0 without a second match, 1 for d2 = 0), f1 sqrt(d1)/512,
// f2..f5 keypoints mapped to [-1, 1], f6 log2 scale ratio (clamped to +-8),
// f7 orientation difference wrapped to [-pi, pi) over pi; rows past count are 0
__global__ void features_kernel(const FeatBatch batch)
{
    const FeatTask& T = batch.t[blockIdx.y];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= T.maxMatches) return;
    const uint32_t cnt = min(*T.count, T.maxMatches);
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (k < cnt) {
        const uint32_t q = T.matches[3 * k], t = T.matches[3 * k + 1];
        const float d1 = __uint_as_float(T.matches[3 * k + 2]);
        const uint32_t d2 = (uint32_t)(T.second[q] >> 32);
        const uint32_t i2 = (uint32_t)T.second[q];
        f[0] = i2 == ~0u ? 0.f : (d2 == 0u ? 1.f : fminf(sqrtf(d1 / (float)d2), 1.f));
        f[1] = sqrtf(d1) * (1.f / 512.f);
        const float4 a = T.kpQ[q], b = T.kpT[t];
        f[2] = a.x * T.invW * 2.f - 1.f;
        f[3] = a.y * T.invH * 2.f - 1.f;
        f[4] = b.x * T.invW * 2.f - 1.f;
        f[5] = b.y * T.invH * 2.f - 1.f;
        f[6] = fminf(fmaxf(log2f(b.z / a.z), -8.f), 8.f);
        float d = b.w - a.w;
        const float pi = 3.14159265358979f;
        d = d - 2.f * pi * floorf((d + pi) * (0.5f / pi));
        f[7] = d * (1.f / pi);
    }
    float4* out = reinterpret_cast<float4*>(T.x + (size_t)k * 8);
    out[0] = make_float4(f[0], f[1], f[2], f[3]);
    out[1] = make_float4(f[4], f[5], f[6], f[7]);
}

static uint32_t grad_grid(uint32_t n)
{
    // enough workgroups to fill the chip, at least one step each
    return std::max(1u, std::min<uint32_t>(512u, div_up(n, kStep)));
}

}  // namespace mlp
}  // namespace rsfm_amd

using namespace rsfm_amd;

extern "C" int rsfm_mlp_param_count(void) { return mlp::kNParam; }

extern "C" int rsfm_mlp_forward(const float* params, const float* x, uint32_t n, float* prob, hipStream_t stream)
{
    RSFM_RETURN_IF(params == nullptr, RSFM_ERR_INVALID, "mlp forward: params is null");
    RSFM_RETURN_IF(n > 0 && (x == nullptr || prob == nullptr), RSFM_ERR_INVALID, "mlp forward: x/prob null");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(x) % 16 != 0, RSFM_ERR_INVALID, "mlp forward: x must be 16-byte aligned");
    if (n == 0) return RSFM_OK;
    const uint32_t grid = std::min<uint32_t>(1024u, div_up(n, mlp::kWaves * 32));
    mlp::forward_kernel<<<grid, 256, 0, stream>>>(params, x, n, prob);
    return check_launch("mlp forward");
}

extern "C" int rsfm_mlp_grad_workspace_size(uint32_t n, size_t* bytes)
{
    RSFM_RETURN_IF(bytes == nullptr, RSFM_ERR_INVALID, "bytes is null");
    *bytes = (size_t)mlp::grad_grid(n) * (mlp::kNParam + 1) * sizeof(float);
    return RSFM_OK;
}

extern "C" int rsfm_mlp_grad(const float* params, const float* x, const float* y, uint32_t n, float* grad, float* loss,
                             void* workspace, size_t workspaceBytes, hipStream_t stream)
{
    RSFM_RETURN_IF(params == nullptr || grad == nullptr, RSFM_ERR_INVALID, "mlp grad: params/grad null");
    RSFM_RETURN_IF(n == 0, RSFM_ERR_INVALID, "mlp grad: empty batch");
    RSFM_RETURN_IF(x == nullptr || y == nullptr, RSFM_ERR_INVALID, "mlp grad: x/y null");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(x) % 16 != 0, RSFM_ERR_INVALID, "mlp grad: x must be 16-byte aligned");
    size_t need = 0;
    rsfm_mlp_grad_workspace_size(n, &need);
    RSFM_RETURN_IF(workspace == nullptr || workspaceBytes < need, RSFM_ERR_INVALID,
                   "mlp grad: workspace too small (" + std::to_string(workspaceBytes) + " < " + std::to_string(need) + ")");
    const uint32_t grid = mlp::grad_grid(n);
    float* part = static_cast<float*>(workspace);
    mlp::grad_kernel<<<grid, 256, 0, stream>>>(params, x, y, n, 1.f / (float)n, part);
    RSFM_CHECK(check_launch("mlp grad"));
    mlp::reduce_kernel<<<div_up(mlp::kNParam + 1, 256), 256, 0, stream>>>(part, grid, n, grad, loss);
    return check_launch("mlp grad reduce");
}

extern "C" int rsfm_mlp_adam_step(float* params, const float* grad, float* m, float* v, uint32_t step, float lr,
                                  float beta1, float beta2, float eps, hipStream_t stream)
{
    RSFM_RETURN_IF(params == nullptr || grad == nullptr || m == nullptr || v == nullptr, RSFM_ERR_INVALID,
                   "mlp adam: null pointer");
    RSFM_RETURN_IF(step == 0, RSFM_ERR_INVALID, "mlp adam: step counts from 1");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    mlp::adam_kernel<<<div_up(mlp::kNParam, 256), 256, 0, stream>>>(params, grad, m, v, lr, beta1, beta2, eps, c1, c2);
    return check_launch("mlp adam");
}

extern "C" int rsfm_proposal_features_batch(const rsfm_feature_task* tasks, size_t nbTasks, hipStream_t stream)
{
    using namespace mlp;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "features: tasks is null");
    for (size_t b = 0; b < nbTasks; b += kFeatMaxTasks) {
        const size_t n = std::min<size_t>(kFeatMaxTasks, nbTasks - b);
        FeatBatch batch{};
        uint32_t maxM = 0;
        for (size_t i = 0; i < n; i++) {
            const rsfm_feature_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.maxMatches > 0 && (t.matches == nullptr || t.count == nullptr || t.second == nullptr ||
                                                t.kpQuery == nullptr || t.kpTrain == nullptr || t.x == nullptr),
                           RSFM_ERR_INVALID, "features: null pointer");
            RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(t.kpQuery) | reinterpret_cast<uintptr_t>(t.kpTrain) |
                            reinterpret_cast<uintptr_t>(t.x)) % 16,
                           RSFM_ERR_INVALID, "features: keypoints and x must be 16-byte aligned");
            RSFM_RETURN_IF(!(t.width > 0.f && t.height > 0.f), RSFM_ERR_INVALID, "features: image size must be > 0");
            batch.t[i] = FeatTask{reinterpret_cast<const uint32_t*>(t.matches), t.count, t.maxMatches,
                                  reinterpret_cast<const uint64_t*>(t.second),
                                  reinterpret_cast<const float4*>(t.kpQuery), reinterpret_cast<const float4*>(t.kpTrain),
                                  1.f / t.width, 1.f / t.height, t.x};
            maxM = std::max(maxM, t.maxMatches);
        }
        if (maxM == 0) continue;
        features_kernel<<<dim3(div_up(maxM, 256), (uint32_t)n), 256, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("proposal features"));
    }
    return RSFM_OK;
}
