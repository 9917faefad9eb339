// hgeom.hpp - FP64 least-squares refits and the plane/parallax geometry of
// DEGENSAC on the GPU, single-lane routines with a fixed operation order
// (sources are built with -ffp-contract=off; IEEE division and square root),
// so the CPU oracle can restate them bit for bit.
//
// Reference semantics:
//  - refit of F on inliers: findEpipolarity (legacy/geometry.hpp:216-259),
//    Hartley normalisation (geometry.hpp:106-135), rank 2 by removing the
//    smallest singular direction;
//  - refit of H on inliers: findHomography (geometry.hpp:139-209), DLT rows
//    [-x, -y, -1, 0, 0, 0, x x', y x', x'] / [0, 0, 0, -x, -y, -1, x y', y y', y'],
//    row-major h, H = T1^-1 H^ T0;
//  - H from F and three correspondences: findHomographyFromEpipolarity
//    (legacy/geometry.cpp:311-327): A = [e']x F, b_i = (x'_i x A x_i).(x'_i x e')
//    / |x'_i x e'|^2, H = A - e' (M^-1 b)^T, M rows = x_i^T;
//  - F from H and two correspondences (plane and parallax):
//    findEpipolarityFromHomography (legacy/geometry.cpp:330-339):
//    l_i = (H x_i) x x'_i, e' = l_0 x l_1, F = [e']x H;
//  - homography inlier test: checkProjectiveTransformation
//    (legacy/geometry.cpp:41-53), expression by expression in FP32.
// Own numerics: the least-squares null vector comes from the FP64 normal
// matrix N = A^T A (entries accumulated point by point in index order), a
// Cholesky factor of N + 1e-10 trace(N) I and five steps of inverse iteration
// from the all-ones vector; the left epipole of F is the largest of the three
// column cross products; the 3x3 solve is Cramer's rule. Every result is
// scaled to unit Frobenius norm and rounded to FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rsfm_amd {
namespace hgeom {

constexpr double kSqrt2 = 1.4142135623730951;
constexpr double kShift = 1e-10;
constexpr int kInverseIters = 5;
constexpr int kJacobiSweepsD = 8;

__device__ inline void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline double dot3(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

__device__ inline double det3(const double G[3][3])
{
    return G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
           G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
}

// cyclic Jacobi on a symmetric 3x3 (FP64 form of fgeom::jacobi3)
__device__ inline void jacobi3d(double S[3][3], double V[3][3])
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < kJacobiSweepsD; sweep++) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int p = r < 2 ? 0 : 1;
            const int q = r == 0 ? 1 : 2;
            const int o = 3 - p - q;
            const double apq = S[p][q];
            if (apq == 0.0) continue;
            const double theta = (S[q][q] - S[p][p]) / (2.0 * apq);
            const double at = fabs(theta);
            double t = 1.0 / (at + sqrt(theta * theta + 1.0));
            if (theta < 0.0) t = -t;
            const double c = 1.0 / sqrt(t * t + 1.0);
            const double s = t * c;
            const double spp = S[p][p] - t * apq;
            const double sqq = S[q][q] + t * apq;
            const double sop = c * S[o][p] - s * S[o][q];
            const double soq = s * S[o][p] + c * S[o][q];
            S[p][p] = spp;
            S[q][q] = sqq;
            S[p][q] = 0.0;
            S[q][p] = 0.0;
            S[o][p] = sop;
            S[p][o] = sop;
            S[o][q] = soq;
            S[q][o] = soq;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
    }
}

// null vector of the symmetric PSD 9x9 N (full storage): Cholesky of
// N + kShift trace(N) I, then inverse iteration; false when N is not usable
__device__ inline bool null_vector9(const double N[9][9], double v[9])
{
    double tr = 0.0;
#pragma unroll
    for (int a = 0; a < 9; a++) tr = tr + N[a][a];
    if (!(tr > 0.0) || !isfinite(tr)) return false;
    const double sigma = kShift * tr;
    double L[9][9], rd[9];  // rd: reciprocals of the diagonal, so every later step multiplies
#pragma unroll
    for (int j = 0; j < 9; j++) {
        double s = N[j][j] + sigma;
#pragma unroll
        for (int k = 0; k < j; k++) s = s - L[j][k] * L[j][k];
        if (!(s > 0.0)) return false;
        L[j][j] = sqrt(s);
        rd[j] = 1.0 / L[j][j];
#pragma unroll
        for (int i = j + 1; i < 9; i++) {
            double t = N[i][j];
#pragma unroll
            for (int k = 0; k < j; k++) t = t - L[i][k] * L[j][k];
            L[i][j] = t * rd[j];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; i++) v[i] = 1.0;
    for (int it = 0; it < kInverseIters; it++) {
        double y[9], z[9];
#pragma unroll
        for (int i = 0; i < 9; i++) {
            double t = v[i];
#pragma unroll
            for (int k = 0; k < i; k++) t = t - L[i][k] * y[k];
            y[i] = t * rd[i];
        }
#pragma unroll
        for (int i = 8; i >= 0; i--) {
            double t = y[i];
#pragma unroll
            for (int k = i + 1; k < 9; k++) t = t - L[k][i] * z[k];
            z[i] = t * rd[i];
        }
        double nrm2 = 0.0;
#pragma unroll
        for (int i = 0; i < 9; i++) nrm2 = nrm2 + z[i] * z[i];
        if (!(nrm2 > 0.0) || !isfinite(nrm2)) return false;
        const double inv = 1.0 / sqrt(nrm2);
#pragma unroll
        for (int i = 0; i < 9; i++) v[i] = z[i] * inv;
    }
    return true;
}

// Hartley normalisation parameters of one refit
struct Norm {
    double c[4];   // centroids x0, y0, x1, y1
    double s0, s1; // scales of image 0 and image 1
};

// unit-Frobenius FP32 copy of a 3x3; false when the norm is zero or not finite
__device__ inline bool to_unit_f32(const double R[3][3], float out[9])
{
    double fro = 0.0;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) fro = fro + R[r][c] * R[r][c];
    fro = sqrt(fro);
    if (!(fro > 0.0) || !isfinite(fro)) return false;
    const double inv = 1.0 / fro;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) out[r * 3 + c] = (float)(R[r][c] * inv);
    return true;
}

// F from the normalised null vector f (column-major vec): rank 2 by removing
// the smallest right singular direction, then F = T1^T F^ T0
__device__ inline bool finish_f(const double f[9], const Norm& nm, float F[9])
{
    double Fh[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) Fh[r][c] = f[3 * c + r];
    double S[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] = (Fh[0][i] * Fh[0][j] + Fh[1][i] * Fh[1][j]) + Fh[2][i] * Fh[2][j];
    jacobi3d(S, V);
    int m = 0;
    if (S[1][1] < S[0][0]) m = 1;
    if (S[2][2] < S[m][m]) m = 2;
    const double v3[3] = {V[0][m], V[1][m], V[2][m]};
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double fv = (Fh[r][0] * v3[0] + Fh[r][1] * v3[1]) + Fh[r][2] * v3[2];
#pragma unroll
        for (int c = 0; c < 3; c++) Fh[r][c] = Fh[r][c] - fv * v3[c];
    }
    const double a0 = -(nm.s0 * nm.c[0]), b0 = -(nm.s0 * nm.c[1]);
    const double a1 = -(nm.s1 * nm.c[2]), b1 = -(nm.s1 * nm.c[3]);
    double G[3][3], R[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        G[r][0] = Fh[r][0] * nm.s0;
        G[r][1] = Fh[r][1] * nm.s0;
        G[r][2] = (Fh[r][0] * a0 + Fh[r][1] * b0) + Fh[r][2];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        R[0][c] = nm.s1 * G[0][c];
        R[1][c] = nm.s1 * G[1][c];
        R[2][c] = (a1 * G[0][c] + b1 * G[1][c]) + G[2][c];
    }
    return to_unit_f32(R, F);
}

// H from the normalised null vector h (row-major): H = T1^-1 H^ T0
__device__ inline bool finish_h(const double h[9], const Norm& nm, float H[9])
{
    const double a0 = -(nm.s0 * nm.c[0]), b0 = -(nm.s0 * nm.c[1]);
    double G[3][3], R[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        G[r][0] = h[3 * r + 0] * nm.s0;
        G[r][1] = h[3 * r + 1] * nm.s0;
        G[r][2] = (h[3 * r + 0] * a0 + h[3 * r + 1] * b0) + h[3 * r + 2];
    }
    const double inv1 = 1.0 / nm.s1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        R[0][c] = G[0][c] * inv1 + nm.c[2] * G[2][c];
        R[1][c] = G[1][c] * inv1 + nm.c[3] * G[2][c];
        R[2][c] = G[2][c];
    }
    return to_unit_f32(R, H);
}

// H compatible with F through three correspondences; false when degenerate
__device__ inline bool h_from_f(const float F[9], const float4 p[3], float H[9])
{
    double Fd[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) Fd[r][c] = (double)F[r * 3 + c];
    const double col[3][3] = {{Fd[0][0], Fd[1][0], Fd[2][0]}, {Fd[0][1], Fd[1][1], Fd[2][1]},
                              {Fd[0][2], Fd[1][2], Fd[2][2]}};
    double cand[3][3];
    cross3(col[0], col[1], cand[0]);
    cross3(col[0], col[2], cand[1]);
    cross3(col[1], col[2], cand[2]);
    int bi = 0;
    double bn = dot3(cand[0], cand[0]);
#pragma unroll
    for (int k = 1; k < 3; k++) {
        const double nk = dot3(cand[k], cand[k]);
        if (nk > bn) {
            bn = nk;
            bi = k;
        }
    }
    if (!(bn > 0.0) || !isfinite(bn)) return false;
    const double inve = 1.0 / sqrt(bn);
    double e[3];
#pragma unroll
    for (int k = 0; k < 3; k++) e[k] = cand[bi][k] * inve;
    double A[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double ac[3];
        cross3(e, col[c], ac);
        A[0][c] = ac[0];
        A[1][c] = ac[1];
        A[2][c] = ac[2];
    }
    double b[3], M[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double x[3] = {(double)p[i].x, (double)p[i].y, 1.0};
        const double xp[3] = {(double)p[i].z, (double)p[i].w, 1.0};
        double Ax[3];
#pragma unroll
        for (int r = 0; r < 3; r++) Ax[r] = (A[r][0] * x[0] + A[r][1] * x[1]) + A[r][2];
        double u[3], w[3];
        cross3(xp, Ax, u);
        cross3(xp, e, w);
        const double ww = dot3(w, w);
        if (!(ww > 0.0)) return false;
        b[i] = dot3(u, w) / ww;
        M[i][0] = x[0];
        M[i][1] = x[1];
        M[i][2] = 1.0;
    }
    const double dM = det3(M);
    if (!(dM != 0.0) || !isfinite(dM)) return false;
    double v[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double Mk[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) Mk[r][c] = (c == k) ? b[r] : M[r][c];
        v[k] = det3(Mk) / dM;
    }
    double R[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) R[r][c] = A[r][c] - e[r] * v[c];
    return to_unit_f32(R, H);
}

// F = [e']x H with e' the intersection of the two parallax lines
__device__ inline bool f_from_h(const float H[9], float4 pa, float4 pb, float F[9])
{
    double Hd[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) Hd[r][c] = (double)H[r * 3 + c];
    double l[2][3];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const float4 q = i == 0 ? pa : pb;
        const double x0 = q.x, y0 = q.y;
        const double xp[3] = {(double)q.z, (double)q.w, 1.0};
        double Hx[3];
#pragma unroll
        for (int r = 0; r < 3; r++) Hx[r] = (Hd[r][0] * x0 + Hd[r][1] * y0) + Hd[r][2];
        cross3(Hx, xp, l[i]);
    }
    double e[3];
    cross3(l[0], l[1], e);
    double R[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double hc[3] = {Hd[0][c], Hd[1][c], Hd[2][c]};
        double fc[3];
        cross3(e, hc, fc);
        R[0][c] = fc[0];
        R[1][c] = fc[1];
        R[2][c] = fc[2];
    }
    return to_unit_f32(R, F);
}

// checkProjectiveTransformation (legacy/geometry.cpp:41-53) for one
// correspondence: |H x - x' z|^2 < t^2 min(z^2, |det H_2x2|)
__device__ inline bool inlier_h(const float H[9], float4 p, float t2)
{
    const float z = (H[8] + H[6] * p.x) + H[7] * p.y;
    const float z2 = z * z;
    const float d = fabsf(H[0] * H[4] - H[3] * H[1]);
    const float sc = d < z2 ? d : z2;
    const float u = ((H[2] + H[0] * p.x) + H[1] * p.y) - p.z * z;
    const float v = ((H[5] + H[3] * p.x) + H[4] * p.y) - p.w * z;
    return u * u + v * v < t2 * sc;
}

}  // namespace hgeom
}  // namespace rsfm_amd
