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

// null vector of the 8 x 9 system whose transpose is M (9 x 8): Householder QR
// of M, f = Q e9 (the last column of Q). M is overwritten.
__device__ inline void null_vector_9x8(float M[9][8], float f[9])
{
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
#pragma unroll
    for (int r = 0; r < 9; r++) f[r] = r == 8 ? 1.f : 0.f;
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
    float f[9];
    null_vector_9x8(M, f);
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

// ---- 4-point homography (findHomography, legacy/geometry.hpp:139-209, the
// hypothesis fit of findHomographyRansacWithInsufficientTiePts<0>,
// ImagePair.cpp:349-378): Hartley normalisation as in fit8, the DLT rows
// [-x, -y, -1, 0, 0, 0, x x', y x', x'] and [0, 0, 0, -x, -y, -1, x y', y y', y']
// of each correspondence, the null vector of the 8 x 9 system (Householder QR
// of its transpose), row-major H^, H = T1^-1 H^ T0, unit Frobenius norm.
__device__ inline bool fit_h4(const float4 pts[4], float H[9])
{
    float c[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        c[0] = c[0] + pts[i].x;
        c[1] = c[1] + pts[i].y;
        c[2] = c[2] + pts[i].z;
        c[3] = c[3] + pts[i].w;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = c[k] / 4.f;
    float n0 = 0.f, n1 = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float d0 = pts[i].x - c[0], d1 = pts[i].y - c[1], d2 = pts[i].z - c[2], d3 = pts[i].w - c[3];
        n0 = n0 + sqrtf(d0 * d0 + d1 * d1);
        n1 = n1 + sqrtf(d2 * d2 + d3 * d3);
    }
    n0 = n0 / 4.f;
    n1 = n1 / 4.f;
    if (!(n0 > 0.f) || !(n1 > 0.f)) return false;
    const float sqrt2 = 1.41421354f;
    const float s0 = sqrt2 / n0, s1 = sqrt2 / n1;
    float M[9][8];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float x = (pts[i].x - c[0]) * s0, y = (pts[i].y - c[1]) * s0;
        const float xp = (pts[i].z - c[2]) * s1, yp = (pts[i].w - c[3]) * s1;
        const int a = 2 * i, b = 2 * i + 1;
        M[0][a] = -x;
        M[1][a] = -y;
        M[2][a] = -1.f;
        M[3][a] = 0.f;
        M[4][a] = 0.f;
        M[5][a] = 0.f;
        M[6][a] = x * xp;
        M[7][a] = y * xp;
        M[8][a] = xp;
        M[0][b] = 0.f;
        M[1][b] = 0.f;
        M[2][b] = 0.f;
        M[3][b] = -x;
        M[4][b] = -y;
        M[5][b] = -1.f;
        M[6][b] = x * yp;
        M[7][b] = y * yp;
        M[8][b] = yp;
    }
    float h[9];
    null_vector_9x8(M, h);
    const float a0 = -(s0 * c[0]), b0 = -(s0 * c[1]);
    float G[3][3], R[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        G[r][0] = h[3 * r + 0] * s0;
        G[r][1] = h[3 * r + 1] * s0;
        G[r][2] = (h[3 * r + 0] * a0 + h[3 * r + 1] * b0) + h[3 * r + 2];
    }
    const float inv1 = 1.f / s1;
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
        R[0][cc] = G[0][cc] * inv1 + c[2] * G[2][cc];
        R[1][cc] = G[1][cc] * inv1 + c[3] * G[2][cc];
        R[2][cc] = G[2][cc];
    }
    float fro = 0.f;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) fro = fro + R[r][cc] * R[r][cc];
    fro = sqrtf(fro);
    if (!(fro > 0.f) || !isfinite(fro)) return false;
    const float invf = 1.f / fro;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) H[r * 3 + cc] = R[r][cc] * invf;
    return true;
}

// ---- 7-point solver (findEpipolarityWith7Points, legacy/geometry.cpp:221-307)
// Own formulation, deterministic bit for bit: Hartley normalisation and a
// Householder QR of the transposed 7x9 system in FP32 give the two null
// vectors f1, f2 (last two columns of Q); the singularity constraint
// det(F2 + t (F1 - F2)) = 0 is a cubic whose coefficients follow exactly from
// four FP64 determinants (t = 0, 1, -1, 2); its real roots are bracketed by the
// critical points and the Cauchy bound and found by 80 bisection steps in
// FP64 (no transcendental functions). Each root gives F = F2 + t (F1 - F2)
// (FP64, rounded to FP32), denormalised as in fit8 and scaled to unit
// Frobenius norm. Up to three solutions.
__device__ inline double det3d(const double G[3][3])
{
    return G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
           G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
}

__device__ inline double cubic_at(const double a[4], double t) { return ((a[3] * t + a[2]) * t + a[1]) * t + a[0]; }

// root of the cubic in [lo, hi] with p(lo), p(hi) of opposite sign (or zero)
__device__ inline double bisect(const double a[4], double lo, double hi)
{
    double plo = cubic_at(a, lo);
    if (plo == 0.0) return lo;
    if (cubic_at(a, hi) == 0.0) return hi;
    for (int it = 0; it < 80; it++) {
        const double mid = 0.5 * (lo + hi);
        const double pm = cubic_at(a, mid);
        if (pm == 0.0) return mid;
        if ((pm < 0.0) == (plo < 0.0)) {
            lo = mid;
            plo = pm;
        } else {
            hi = mid;
        }
    }
    return 0.5 * (lo + hi);
}

// real roots of a[3] t^3 + a[2] t^2 + a[1] t + a[0] (a[3] != 0), ascending
__device__ inline int cubic_roots(const double a[4], double r[3])
{
    const double inv = 1.0 / fabs(a[3]);
    double B = fabs(a[2]) * inv;
    B = fabs(a[1]) * inv > B ? fabs(a[1]) * inv : B;
    B = fabs(a[0]) * inv > B ? fabs(a[0]) * inv : B;
    B = B + 1.0;
    double e[4] = {-B, 0.0, 0.0, B};
    int ne = 2;
    const double disc = a[2] * a[2] - 3.0 * a[3] * a[1];
    if (disc > 0.0) {
        const double sq = sqrt(disc);
        double c1 = (-a[2] - sq) / (3.0 * a[3]), c2 = (-a[2] + sq) / (3.0 * a[3]);
        if (c2 < c1) {
            const double tmp = c1;
            c1 = c2;
            c2 = tmp;
        }
        e[1] = c1;
        e[2] = c2;
        e[3] = B;
        ne = 4;
    } else {
        e[1] = B;
    }
    int n = 0;
    for (int i = 0; i + 1 < ne; i++) {
        const double lo = e[i], hi = e[i + 1];
        const double plo = cubic_at(a, lo), phi = cubic_at(a, hi);
        if ((plo < 0.0 && phi < 0.0) || (plo > 0.0 && phi > 0.0)) continue;
        const double x = bisect(a, lo, hi);
        if (n > 0 && x == r[n - 1]) continue;
        r[n++] = x;
    }
    return n;
}

// pts[i] = (x0, y0, x1, y1); Fs[k] row-major, unit Frobenius; returns the count
__device__ inline int fit7(const float4 pts[7], float Fs[3][9])
{
    float c[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 7; i++) {
        c[0] = c[0] + pts[i].x;
        c[1] = c[1] + pts[i].y;
        c[2] = c[2] + pts[i].z;
        c[3] = c[3] + pts[i].w;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = c[k] / 7.f;
    float n0 = 0.f, n1 = 0.f;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const float d0 = pts[i].x - c[0], d1 = pts[i].y - c[1], d2 = pts[i].z - c[2], d3 = pts[i].w - c[3];
        n0 = n0 + sqrtf(d0 * d0 + d1 * d1);
        n1 = n1 + sqrtf(d2 * d2 + d3 * d3);
    }
    n0 = n0 / 7.f;
    n1 = n1 / 7.f;
    if (!(n0 > 0.f) || !(n1 > 0.f)) return 0;
    const float sqrt2 = 1.41421354f;
    const float s0 = sqrt2 / n0, s1 = sqrt2 / n1;
    float M[9][7];
#pragma unroll
    for (int i = 0; i < 7; i++) {
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
    float vkk[7], beta[7];
#pragma unroll
    for (int k = 0; k < 7; k++) {
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
        for (int j = k + 1; j < 7; j++) {
            float w = vkk[k] * M[k][j];
#pragma unroll
            for (int r = k + 1; r < 9; r++) w = w + M[r][k] * M[r][j];
            const float bw = beta[k] * w;
            M[k][j] = M[k][j] - bw * vkk[k];
#pragma unroll
            for (int r = k + 1; r < 9; r++) M[r][j] = M[r][j] - bw * M[r][k];
        }
    }
    // f1 = Q e8, f2 = Q e7
    float f[2][9];
#pragma unroll
    for (int q = 0; q < 2; q++) {
#pragma unroll
        for (int r = 0; r < 9; r++) f[q][r] = (r == 8 - q) ? 1.f : 0.f;
#pragma unroll
        for (int k = 6; k >= 0; k--) {
            if (beta[k] == 0.f) continue;
            float w = vkk[k] * f[q][k];
#pragma unroll
            for (int r = k + 1; r < 9; r++) w = w + M[r][k] * f[q][r];
            const float bw = beta[k] * w;
            f[q][k] = f[q][k] - bw * vkk[k];
#pragma unroll
            for (int r = k + 1; r < 9; r++) f[q][r] = f[q][r] - bw * M[r][k];
        }
    }
    // G(t) = F2 + t D with F1 = mat(f[0]), F2 = mat(f[1]) (column-major vec)
    double F2[3][3], D[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
            F2[r][cc] = (double)f[1][3 * cc + r];
            D[r][cc] = (double)f[0][3 * cc + r] - F2[r][cc];
        }
    double d[4];
    const double ts[4] = {0.0, 1.0, -1.0, 2.0};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double G[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) G[r][cc] = F2[r][cc] + ts[q] * D[r][cc];
        d[q] = det3d(G);
    }
    double a[4];
    a[0] = d[0];
    a[2] = 0.5 * (d[1] + d[2]) - d[0];
    const double sOdd = 0.5 * (d[1] - d[2]);
    a[3] = ((d[3] - d[0]) - 4.0 * a[2] - 2.0 * sOdd) / 6.0;
    a[1] = sOdd - a[3];
    double scale = fabs(a[0]);
    for (int q = 1; q < 4; q++) scale = fabs(a[q]) > scale ? fabs(a[q]) : scale;
    if (!(scale > 0.0) || !(fabs(a[3]) > 1e-12 * scale)) return 0;
    double roots[3];
    const int nr = cubic_roots(a, roots);
    const float a0 = -(s0 * c[0]), b0 = -(s0 * c[1]);
    const float a1 = -(s1 * c[2]), b1 = -(s1 * c[3]);
    int out = 0;
    for (int k = 0; k < nr; k++) {
        float Fh[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) Fh[r][cc] = (float)(F2[r][cc] + roots[k] * D[r][cc]);
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
        if (!(fro > 0.f) || !isfinite(fro)) continue;
        const float invf = 1.f / fro;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) Fs[out][r * 3 + cc] = R[r][cc] * invf;
        out++;
    }
    return out;
}

// Inlier tests. kMode 0 = checkEpipolarTransformation (legacy/geometry.cpp:
// 55-86), expression by expression without fused multiply-add: both
// point-to-line distances below t. kMode 1 = Sampson distance below t, this
// library's own test (no reference counterpart), written with fused
// multiply-adds (fmaf: one rounding each, identical on CPU and GPU):
// e = x1^T F x0, e^2 < t^2 (|F^T x1|_12^2 + |F x0|_12^2).
template <int kMode>
__device__ inline bool inlier(const float F[9], float4 p, float t2)
{
    const float x0 = p.x, y0 = p.y, x1 = p.z, y1 = p.w;
    if constexpr (kMode == 0) {
        const float l0a = x1 * F[0] + (y1 * F[3] + F[6]);
        const float l0b = x1 * F[1] + (y1 * F[4] + F[7]);
        const float l0c = x1 * F[2] + (y1 * F[5] + F[8]);
        const float l1a = x0 * F[0] + (y0 * F[1] + F[2]);
        const float l1b = x0 * F[3] + (y0 * F[4] + F[5]);
        const float e = l0a * x0 + (l0b * y0 + l0c);
        const float e2 = e * e;
        const float n0 = l0a * l0a + l0b * l0b;
        const float n1 = l1a * l1a + l1b * l1b;
        return (e2 < t2 * n0) && (e2 < t2 * n1);
    } else {
        const float l0a = fmaf(x1, F[0], fmaf(y1, F[3], F[6]));
        const float l0b = fmaf(x1, F[1], fmaf(y1, F[4], F[7]));
        const float l0c = fmaf(x1, F[2], fmaf(y1, F[5], F[8]));
        const float l1a = fmaf(x0, F[0], fmaf(y0, F[1], F[2]));
        const float l1b = fmaf(x0, F[3], fmaf(y0, F[4], F[5]));
        const float e = fmaf(l0a, x0, fmaf(l0b, y0, l0c));
        const float n = fmaf(l0a, l0a, fmaf(l0b, l0b, fmaf(l1a, l1a, l1b * l1b)));
        return e * e < t2 * n;
    }
}

}  // namespace fgeom
}  // namespace rsfm_amd
