/*
 * pose.c - relative pose from F or H restated on the CPU.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * Reference: decomposeEpipolarity / decomposeEssentialMatrix
 * (legacy/geometry.hpp:268-346), triangulate (legacy/geometry.cpp:146-206),
 * isInFrontOfCam and disambiguateSolutions (ImagePair.cpp:760-786).
 * Repeats rsfm_pose_from_f operation for operation (FP64 E and SVD via a
 * cyclic 3x3 Jacobi on E^T E, u_k = E v_k / |E v_k|, u3 = u1 x u2,
 * v3 = v1 x v2; FP32 triangulation with the adjugate solve). Build with
 * -ffp-contract=off.
 * Pinned in tests by noise-free geometry: one candidate is the true (R, t/|t|)
 * and has every point in front, the four rotations are orthonormal with det +1,
 * [t]x R is proportional to K1^T F K0; the triangulated points of noise-free
 * correspondences are the true 3D points. The reference's Eigen FP32 SVD is not
 * reproduced bit for bit (Eigen is absent here).
 *
 * Homography branch: decomposeHomography (legacy/geometry.hpp:364-447) and
 * solveByHomography's transform (ImagePair.cpp:805-816), repeating
 * rsfm_pose_from_h: FP64 H = K1^-1 G K0, Jacobi on H^T H, eigenvalues over the
 * middle one, det V = +1, u = a v1 +- b v3, R = W U^T, N = v2 x u,
 * T/d = (H - R) N for H and -H, kept where N . T/d < 1. Pinned in tests by
 * noise-free plane scenes: every candidate satisfies H = R + (T/d) N^T up to
 * the sign and scale of H with R orthonormal, det +1 and |N| = 1; one candidate
 * is the true (R, t/|t|, plane normal) and has every point in front.
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <string.h>

static void jacobi3_pose(double S[3][3], double V[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 8; sweep++) {
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
            for (int k = 0; k < 3; k++) {
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
    }
}

static void cross3d(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

int or_pose_decompose(const float Fin[9], const float c0[4], const float c1[4], float out[48])
{
    double F[3][3], G[3][3], E[3][3], S[3][3], V[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) F[r][c] = (double)Fin[r * 3 + c];
    const double fx0 = c0[0], fy0 = c0[1], cx0 = c0[2], cy0 = c0[3];
    const double fx1 = c1[0], fy1 = c1[1], cx1 = c1[2], cy1 = c1[3];
    for (int a = 0; a < 3; a++) {
        G[a][0] = F[a][0] * fx0;
        G[a][1] = F[a][1] * fy0;
        G[a][2] = (F[a][0] * cx0 + F[a][1] * cy0) + F[a][2];
    }
    for (int c = 0; c < 3; c++) {
        E[0][c] = fx1 * G[0][c];
        E[1][c] = fy1 * G[1][c];
        E[2][c] = (cx1 * G[0][c] + cy1 * G[1][c]) + G[2][c];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i][j] = (E[0][i] * E[0][j] + E[1][i] * E[1][j]) + E[2][i] * E[2][j];
    jacobi3_pose(S, V);
    int i0 = 0;
    if (S[1][1] > S[i0][i0]) i0 = 1;
    if (S[2][2] > S[i0][i0]) i0 = 2;
    int i1 = i0 == 0 ? 1 : 0;
    for (int k = 0; k < 3; k++)
        if (k != i0 && k != i1 && S[k][k] > S[i1][i1]) i1 = k;
    const double v1[3] = {V[0][i0], V[1][i0], V[2][i0]};
    const double v2[3] = {V[0][i1], V[1][i1], V[2][i1]};
    double u1[3], u2[3], u3[3], v3[3];
    for (int r = 0; r < 3; r++) {
        u1[r] = (E[r][0] * v1[0] + E[r][1] * v1[1]) + E[r][2] * v1[2];
        u2[r] = (E[r][0] * v2[0] + E[r][1] * v2[1]) + E[r][2] * v2[2];
    }
    const double n1 = sqrt((u1[0] * u1[0] + u1[1] * u1[1]) + u1[2] * u1[2]);
    const double n2 = sqrt((u2[0] * u2[0] + u2[1] * u2[1]) + u2[2] * u2[2]);
    if (!(n1 > 0.0) || !(n2 > 0.0) || !isfinite(n1) || !isfinite(n2)) return 0;
    for (int r = 0; r < 3; r++) {
        u1[r] = u1[r] / n1;
        u2[r] = u2[r] / n2;
    }
    cross3d(u1, u2, u3);
    cross3d(v1, v2, v3);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            const double a = u1[r] * v2[c] - u2[r] * v1[c];
            const double b = u3[r] * v3[c];
            const float R0 = (float)(a + b), R1 = (float)(b - a);
            out[0 * 12 + r * 3 + c] = R0;
            out[1 * 12 + r * 3 + c] = R0;
            out[2 * 12 + r * 3 + c] = R1;
            out[3 * 12 + r * 3 + c] = R1;
        }
    for (int r = 0; r < 3; r++) {
        const float t = (float)u3[r];
        out[0 * 12 + 9 + r] = t;
        out[1 * 12 + 9 + r] = -t;
        out[2 * 12 + 9 + r] = t;
        out[3 * 12 + 9 + r] = -t;
    }
    return 1;
}

/* triangulated point (X) of one correspondence; returns in-front-of-both */
int or_pose_triangulate(const float P[12], const float* p, const float inv0[4], const float inv1[4], float X[3])
{
    const float r0[3] = {inv0[0] * p[0] + inv0[2], inv0[1] * p[1] + inv0[3], 1.f};
    const float q1[3] = {inv1[0] * p[2] + inv1[2], inv1[1] * p[3] + inv1[3], 1.f};
    float r1[3], C1[3];
    for (int k = 0; k < 3; k++) {
        r1[k] = (P[0 * 3 + k] * q1[0] + P[1 * 3 + k] * q1[1]) + P[2 * 3 + k] * q1[2];
        C1[k] = -((P[0 * 3 + k] * P[9] + P[1 * 3 + k] * P[10]) + P[2 * 3 + k] * P[11]);
    }
    const float l0 = sqrtf((r0[0] * r0[0] + r0[1] * r0[1]) + r0[2] * r0[2]);
    const float l1 = sqrtf((r1[0] * r1[0] + r1[1] * r1[1]) + r1[2] * r1[2]);
    float a[3], b[3];
    for (int k = 0; k < 3; k++) {
        a[k] = r0[k] / l0;
        b[k] = r1[k] / l1;
    }
    float A[3][3], M1[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const float d = i == j ? 1.f : 0.f;
            M1[i][j] = b[i] * b[j] - d;
            A[i][j] = (a[i] * a[j] - d) + M1[i][j];
        }
    float rhs[3];
    for (int i = 0; i < 3; i++) rhs[i] = (M1[i][0] * C1[0] + M1[i][1] * C1[1]) + M1[i][2] * C1[2];
    float adj[3][3];
    adj[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    adj[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    adj[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    adj[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    adj[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    adj[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    adj[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    adj[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    adj[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const float det = (A[0][0] * adj[0][0] + A[0][1] * adj[1][0]) + A[0][2] * adj[2][0];
    const float inv = 1.f / det;
    for (int i = 0; i < 3; i++) X[i] = ((adj[i][0] * rhs[0] + adj[i][1] * rhs[1]) + adj[i][2] * rhs[2]) * inv;
    const float z1 = ((P[6] * X[0] + P[7] * X[1]) + P[8] * X[2]) + P[11];
    return X[2] > 0.f && z1 > 0.f;
}

void or_pose_from_f(const float F[9], const float c0[4], const float c1[4], const float* pts, uint32_t n,
                    const uint8_t* mask, float ratio, float poses[48], uint32_t result[10])
{
    float P[48];
    const int ok = or_pose_decompose(F, c0, c1, P);
    float inv0[4], inv1[4];
    inv0[0] = 1.f / c0[0];
    inv0[1] = 1.f / c0[1];
    inv0[2] = -c0[2] * inv0[0];
    inv0[3] = -c0[3] * inv0[1];
    inv1[0] = 1.f / c1[0];
    inv1[1] = 1.f / c1[1];
    inv1[2] = -c1[2] * inv1[0];
    inv1[3] = -c1[3] * inv1[1];
    uint32_t cnt[4] = {0, 0, 0, 0}, in = 0;
    if (ok)
        for (uint32_t i = 0; i < n; i++) {
            if (mask && !mask[i]) continue;
            in++;
            for (int k = 0; k < 4; k++) {
                float X[3];
                cnt[k] += (uint32_t)or_pose_triangulate(P + 12 * k, pts + (size_t)i * 4, inv0, inv1, X);
            }
        }
    uint32_t order[4] = {0, 1, 2, 3};
    for (int a = 1; a < 4; a++)
        for (int b = a; b > 0 && cnt[order[b]] > cnt[order[b - 1]]; b--) {
            const uint32_t t = order[b];
            order[b] = order[b - 1];
            order[b - 1] = t;
        }
    const uint32_t min_cnt = (uint32_t)roundf((float)in * ratio);
    uint32_t kept = 0;
    while (kept < 4 && ok && cnt[order[kept]] >= min_cnt && cnt[order[kept]] >= 2) kept++;
    for (int k = 0; k < 4; k++) {
        result[k] = cnt[k];
        result[4 + k] = order[k];
    }
    result[8] = kept;
    result[9] = in;
    for (int q = 0; q < 48; q++) poses[q] = ok ? P[q] : 0.f;
}

static void hs_scaled(const double Hn[3][3], double sc, int neg, double Hs[3][3])
{
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Hs[r][c] = neg ? -(Hn[r][c] / sc) : Hn[r][c] / sc;
}

static double dot3d(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

int or_pose_decompose_h(const float Hin[9], const float c0[4], const float c1[4], float out[96], float nrm[24])
{
    double G[3][3], Hn[3][3], S[3][3], V[3][3];
    const double fx0 = c0[0], fy0 = c0[1], cx0 = c0[2], cy0 = c0[3];
    const double fx1 = c1[0], fy1 = c1[1], cx1 = c1[2], cy1 = c1[3];
    for (int a = 0; a < 3; a++) {
        const double h0 = Hin[a * 3 + 0], h1 = Hin[a * 3 + 1], h2 = Hin[a * 3 + 2];
        G[a][0] = h0 * fx0;
        G[a][1] = h1 * fy0;
        G[a][2] = (h0 * cx0 + h1 * cy0) + h2;
    }
    for (int c = 0; c < 3; c++) {
        Hn[0][c] = (G[0][c] - cx1 * G[2][c]) / fx1;
        Hn[1][c] = (G[1][c] - cy1 * G[2][c]) / fy1;
        Hn[2][c] = G[2][c];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i][j] = (Hn[0][i] * Hn[0][j] + Hn[1][i] * Hn[1][j]) + Hn[2][i] * Hn[2][j];
    jacobi3_pose(S, V);
    int i0 = 0;
    if (S[1][1] > S[i0][i0]) i0 = 1;
    if (S[2][2] > S[i0][i0]) i0 = 2;
    int i1 = i0 == 0 ? 1 : 0;
    for (int k = 0; k < 3; k++)
        if (k != i0 && k != i1 && S[k][k] > S[i1][i1]) i1 = k;
    const int i2 = 3 - i0 - i1;
    const double lam1 = S[i1][i1];
    if (!(lam1 > 0.0) || !isfinite(S[i0][i0])) return 0;
    const double sc = sqrt(lam1);
    const double s0 = S[i0][i0] / lam1, s2 = S[i2][i2] / lam1;
    double v1[3] = {V[0][i0], V[1][i0], V[2][i0]};
    double v2[3] = {V[0][i1], V[1][i1], V[2][i1]};
    double v3[3] = {V[0][i2], V[1][i2], V[2][i2]};
    double x[3];
    cross3d(v2, v3, x);
    if (dot3d(v1, x) < 0.0)
        for (int r = 0; r < 3; r++) {
            v1[r] = -v1[r];
            v2[r] = -v2[r];
            v3[r] = -v3[r];
        }
    const double dd = s0 - s2;
    if (!(dd > 0.0)) return 0;
    const double cc = 1.0 / sqrt(dd);
    const double a = sqrt(1.0 - s2) * cc, b = sqrt(s0 - 1.0) * cc;
    double u[2][3], nv[2][3];
    for (int r = 0; r < 3; r++) {
        u[0][r] = a * v1[r] + b * v3[r];
        u[1][r] = a * v1[r] - b * v3[r];
    }
    cross3d(v2, u[0], nv[0]);
    cross3d(v2, u[1], nv[1]);
    int nOut = 0;
    for (int sgn = 0; sgn < 2; sgn++) {
        double Hs[3][3], w0[3], R[2][3][3], T[2][3];
        hs_scaled(Hn, sc, sgn, Hs);
        for (int r = 0; r < 3; r++) w0[r] = (Hs[r][0] * v2[0] + Hs[r][1] * v2[1]) + Hs[r][2] * v2[2];
        for (int k = 0; k < 2; k++) {
            double w1[3], w2[3];
            for (int r = 0; r < 3; r++) w1[r] = (Hs[r][0] * u[k][0] + Hs[r][1] * u[k][1]) + Hs[r][2] * u[k][2];
            cross3d(w0, w1, w2);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) R[k][r][c] = (w0[r] * v2[c] + w1[r] * u[k][c]) + w2[r] * nv[k][c];
            for (int r = 0; r < 3; r++)
                T[k][r] = ((Hs[r][0] - R[k][r][0]) * nv[k][0] + (Hs[r][1] - R[k][r][1]) * nv[k][1]) +
                          (Hs[r][2] - R[k][r][2]) * nv[k][2];
        }
        for (int q = 0; q < 4; q++) {
            const int k = q & 1;
            if (!(dot3d(nv[k], T[k]) < 1.0)) continue;
            const double tn = sqrt(dot3d(T[k], T[k]));
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) out[nOut * 12 + r * 3 + c] = (float)R[k][r][c];
            for (int r = 0; r < 3; r++) {
                const double t = tn > 0.0 ? T[k][r] / tn : T[k][r];
                out[nOut * 12 + 9 + r] = (float)(q < 2 ? t : -t);
                nrm[nOut * 3 + r] = (float)(q < 2 ? nv[k][r] : -nv[k][r]);
            }
            nOut++;
        }
    }
    return nOut;
}

void or_pose_from_h(const float H[9], const float c0[4], const float c1[4], const float* pts, uint32_t n,
                    const uint8_t* mask, float ratio, float poses[96], float normals[24], uint32_t result[19])
{
    float P[96], N[24];
    memset(P, 0, sizeof(P));
    memset(N, 0, sizeof(N));
    const int nc = or_pose_decompose_h(H, c0, c1, P, N);
    float inv0[4], inv1[4];
    inv0[0] = 1.f / c0[0];
    inv0[1] = 1.f / c0[1];
    inv0[2] = -c0[2] * inv0[0];
    inv0[3] = -c0[3] * inv0[1];
    inv1[0] = 1.f / c1[0];
    inv1[1] = 1.f / c1[1];
    inv1[2] = -c1[2] * inv1[0];
    inv1[3] = -c1[3] * inv1[1];
    uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0}, in = 0;
    if (nc > 0)
        for (uint32_t i = 0; i < n; i++) {
            if (mask && !mask[i]) continue;
            in++;
            for (int k = 0; k < nc; k++) {
                float X[3];
                cnt[k] += (uint32_t)or_pose_triangulate(P + 12 * k, pts + (size_t)i * 4, inv0, inv1, X);
            }
        }
    uint32_t order[8];
    for (int k = 0; k < 8; k++) order[k] = (uint32_t)k;
    for (int a = 1; a < nc; a++)
        for (int b = a; b > 0 && cnt[order[b]] > cnt[order[b - 1]]; b--) {
            const uint32_t t = order[b];
            order[b] = order[b - 1];
            order[b - 1] = t;
        }
    const uint32_t min_cnt = (uint32_t)roundf((float)in * ratio);
    uint32_t kept = 0;
    while ((int)kept < nc && cnt[order[kept]] >= min_cnt && cnt[order[kept]] >= 2) kept++;
    for (int k = 0; k < 8; k++) {
        result[k] = cnt[k];
        result[8 + k] = order[k];
    }
    result[16] = kept;
    result[17] = in;
    result[18] = (uint32_t)nc;
    memcpy(poses, P, sizeof(P));
    if (normals) memcpy(normals, N, sizeof(N));
}
