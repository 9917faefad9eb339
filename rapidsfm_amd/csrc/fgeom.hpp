// fgeom.hpp - FP32 two-view geometry on the GPU: 8-point fundamental matrix
// and the epipolar inlier tests.
//
// The operation order is fixed (no fused multiply-add: every source is built
// with -ffp-contract=off; IEEE division and square root) so results are
// reproducible bit for bit; the CPU oracle restates the same sequence.
// Reference semantics: findEpipolarity (legacy/geometry.hpp:216-259) with
// Hartley normalisation (geometry.hpp:106-135): design row
// [x1x0, y1x0, x0, x1y0, y1y0, y0, x1, y1, 1] = column-major vec(F), null
// vector of the 8x9 system, rank 2 by removing the smallest singular
// direction, F = T1^T F^ T0. Here the null vector comes from a Householder QR
// of the transposed system (last column of Q) and the rank-2 projection from
// a cyclic Jacobi eigensolver on F^T F (F' = F - (F v3) v3^T); the result is
// scaled to unit Frobenius norm.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rsfm_amd {
namespace fgeom {

constexpr int kJacobiSweeps = 6;

__device__ inline void jacobi3(float S[3][3], float V[3][3])
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.f : 0.f;
#pragma unroll
    for (int sweep = 0; sweep < kJacobiSweeps; sweep++) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int p = r < 2 ? 0 : 1;
            const int q = r == 0 ? 1 : 2;
            const int o = 3 - p - q;
            const float apq = S[p][q];
            if (apq == 0.f) continue;
            const float theta = (S[q][q] - S[p][p]) / (2.f * apq);
            const float at = fabsf(theta);
            float t = 1.f / (at + sqrtf(theta * theta + 1.f));
            if (theta < 0.f) t = -t;
            const float c = 1.f / sqrtf(t * t + 1.f);
            const float s = t * c;
            const float spp = S[p][p] - t * apq;
            const float sqq = S[q][q] + t * apq;
            const float sop = c * S[o][p] - s * S[o][q];
            const float soq = s * S[o][p] + c * S[o][q];
            S[p][p] = spp;
            S[q][q] = sqq;
            S[p][q] = 0.f;
            S[q][p] = 0.f;
            S[o][p] = sop;
            S[p][o] = sop;
            S[o][q] = soq;
            S[q][o] = soq;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
    }
}

// pts[i] = (x0, y0, x1, y1) of sample i; F row-major; returns false when degenerate
__device__ inline bool fit8(const float4 pts[8], float F[9])
{
    float c[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 8; i++) {
        c[0] = c[0] + pts[i].x;
        c[1] = c[1] + pts[i].y;
        c[2] = c[2] + pts[i].z;
        c[3] = c[3] + pts[i].w;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = c[k] / 8.f;
    float n0 = 0.f, n1 = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float d0 = pts[i].x - c[0], d1 = pts[i].y - c[1], d2 = pts[i].z - c[2], d3 = pts[i].w - c[3];
        n0 = n0 + sqrtf(d0 * d0 + d1 * d1);
        n1 = n1 + sqrtf(d2 * d2 + d3 * d3);
    }
    n0 = n0 / 8.f;
    n1 = n1 / 8.f;
    if (!(n0 > 0.f) || !(n1 > 0.f)) return false;
    const float sqrt2 = 1.41421354f;
    const float s0 = sqrt2 / n0, s1 = sqrt2 / n1;
    // M = P^T (9 x 8); the Householder vector k lives below the diagonal of
    // column k, its head in vkk[k]
    float M[9][8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float x0 = (pts[i].x - c[0]) * s0, y0 = (pts[i].y - c[1]) * s0;
        const float x1 = (pts[i].z - c[2]) * s1, y1 = (pts[i].w - c[3]) * s1;
        M[0][i] = x1 * x0;
        M[1][i] = y1 * x0;
        M[2][i] = x0;
        M[3][i] = x1 * y0;
        M[4][i] = y1 * y0;
        M[5][i] = y0;
        M[6][i] = x1;
        M[7][i] = y1;
        M[8][i] = 1.f;
    }
    float vkk[8], beta[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        float nrm2 = 0.f;
#pragma unroll
        for (int r = k; r < 9; r++) nrm2 = nrm2 + M[r][k] * M[r][k];
        const float nrm = sqrtf(nrm2);
        vkk[k] = 0.f;
        beta[k] = 0.f;
        if (nrm == 0.f) {
#pragma unroll
            for (int r = k + 1; r < 9; r++) M[r][k] = 0.f;
            continue;
        }
        const float alpha = (M[k][k] > 0.f) ? -nrm : nrm;
        vkk[k] = M[k][k] - alpha;
        float vtv = vkk[k] * vkk[k];
#pragma unroll
        for (int r = k + 1; r < 9; r++) vtv = vtv + M[r][k] * M[r][k];
        if (vtv == 0.f) continue;
        beta[k] = 2.f / vtv;
#pragma unroll
        for (int j = k + 1; j < 8; j++) {
            float w = vkk[k] * M[k][j];
#pragma unroll
            for (int r = k + 1; r < 9; r++) w = w + M[r][k] * M[r][j];
            const float bw = beta[k] * w;
            M[k][j] = M[k][j] - bw * vkk[k];
#pragma unroll
            for (int r = k + 1; r < 9; r++) M[r][j] = M[r][j] - bw * M[r][k];
        }
    }
    float f[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f};
#pragma unroll
    for (int k = 7; k >= 0; k--) {
        if (beta[k] == 0.f) continue;
        float w = vkk[k] * f[k];
#pragma unroll
        for (int r = k + 1; r < 9; r++) w = w + M[r][k] * f[r];
        const float bw = beta[k] * w;
        f[k] = f[k] - bw * vkk[k];
#pragma unroll
        for (int r = k + 1; r < 9; r++) f[r] = f[r] - bw * M[r][k];
    }
    float Fh[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) Fh[r][cc] = f[3 * cc + r];
    float S[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] = (Fh[0][i] * Fh[0][j] + Fh[1][i] * Fh[1][j]) + Fh[2][i] * Fh[2][j];
    jacobi3(S, V);
    int m = 0;
    if (S[1][1] < S[0][0]) m = 1;
    if (S[2][2] < S[m][m]) m = 2;
    const float v3[3] = {V[0][m], V[1][m], V[2][m]};
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float fv = (Fh[r][0] * v3[0] + Fh[r][1] * v3[1]) + Fh[r][2] * v3[2];
#pragma unroll
        for (int cc = 0; cc < 3; cc++) Fh[r][cc] = Fh[r][cc] - fv * v3[cc];
    }
    const float a0 = -(s0 * c[0]), b0 = -(s0 * c[1]);
    const float a1 = -(s1 * c[2]), b1 = -(s1 * c[3]);
    float G[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        G[r][0] = Fh[r][0] * s0;
        G[r][1] = Fh[r][1] * s0;
        G[r][2] = (Fh[r][0] * a0 + Fh[r][1] * b0) + Fh[r][2];
    }
    float R[3][3];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
        R[0][cc] = s1 * G[0][cc];
        R[1][cc] = s1 * G[1][cc];
        R[2][cc] = (a1 * G[0][cc] + b1 * G[1][cc]) + G[2][cc];
    }
    float fro = 0.f;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) fro = fro + R[r][cc] * R[r][cc];
    fro = sqrtf(fro);
    if (!(fro > 0.f) || !isfinite(fro)) return false;
    const float inv = 1.f / fro;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) F[r * 3 + cc] = R[r][cc] * inv;
    return true;
}

// checkEpipolarTransformation (legacy/geometry.cpp:55-86), expression by
// expression; kMode 0 = both point-to-line distances below t (reference),
// kMode 1 = Sampson distance below t
template <int kMode>
__device__ inline bool inlier(const float F[9], float4 p, float t2)
{
    const float x0 = p.x, y0 = p.y, x1 = p.z, y1 = p.w;
    const float l0a = x1 * F[0] + (y1 * F[3] + F[6]);
    const float l0b = x1 * F[1] + (y1 * F[4] + F[7]);
    const float l0c = x1 * F[2] + (y1 * F[5] + F[8]);
    const float l1a = x0 * F[0] + (y0 * F[1] + F[2]);
    const float l1b = x0 * F[3] + (y0 * F[4] + F[5]);
    const float e = l0a * x0 + (l0b * y0 + l0c);
    const float e2 = e * e;
    const float n0 = l0a * l0a + l0b * l0b;
    const float n1 = l1a * l1a + l1b * l1b;
    if constexpr (kMode == 0) return (e2 < t2 * n0) && (e2 < t2 * n1);
    else return e2 < t2 * (n0 + n1);
}

}  // namespace fgeom
}  // namespace rsfm_amd
