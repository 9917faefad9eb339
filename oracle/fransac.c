/*
 * fransac.c - fundamental-matrix RANSAC restated on the CPU.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * Reference path: findEpipolarityRansacWithInsufficientTiePts<0> (ImagePair.cpp:148-204)
 * driving ransac<F,8> (ransac.hpp:115-163) with the TableSampler of ransac.hpp:52-82;
 * per hypothesis: sample sanity gate (ImagePair.cpp:160-177), findEpipolarity
 * (legacy/geometry.hpp:216-259: Hartley normalisation geometry.hpp:106-135, 8x9
 * design row [x1x0, y1x0, x0, x1y0, y1y0, y0, x1, y1, 1] = column-major vec(F),
 * null vector, rank 2 by zeroing the smallest singular value, F = T1^T F^ T0),
 * then checkEpipolarTransformation (legacy/geometry.cpp:55-86).
 *
 * Two fits are restated:
 *  - or_f8_fit_f32: FP32 with a fixed operation order (no fused multiply-add,
 *    build with -ffp-contract=off) that the GPU kernel repeats step for step,
 *    so GPU and CPU inlier sets are identical for the same sample indices.
 *    The null vector comes from a Householder QR of P^T (9x8, last column of
 *    Q) and the rank-2 projection from a cyclic Jacobi eigensolver on F^T F
 *    (6 sweeps): F' = F - (F v3) v3^T, v3 = smallest eigenvector. This is the
 *    same mathematics as the reference's JacobiSVD route, not its FP32 bits
 *    (Eigen, an unpinned system package, is absent here).
 *  - or_f8_fit_f64: FP64 model used as the accuracy oracle (Jacobi SVD of P
 *    via the eigen-decomposition of P^T P to convergence, Jacobi on F^T F).
 * Pinned in tests by: noise-free two-view geometry (the fit must recover the
 * closed-form F = K^-T [t]x R K^-1 up to scale), numpy SVD null vectors, and
 * hand-solved epipolar distances for rectified pairs.
 * The two-sided inlier test (mode 0) follows geometry.cpp:66-83 expression by
 * expression without contraction; the Sampson mode (1) is this library's own
 * test, e^2 < t^2 (l0x^2+l0y^2+l1x^2+l1y^2), written with fmaf exactly where
 * the GPU kernel fuses (one rounding per fmaf on both sides).
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define JACOBI3_SWEEPS 6

/* ---------------- FP32 fit (operation order shared with the GPU kernel) ---------------- */

static void jacobi3_f32(float S[3][3], float V[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.f : 0.f;
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < JACOBI3_SWEEPS; sweep++)
        for (int r = 0; r < 3; r++) {
            const int p = P[r], q = Q[r], o = 3 - p - q;
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
            S[p][p] = spp; S[q][q] = sqq;
            S[p][q] = 0.f; S[q][p] = 0.f;
            S[o][p] = sop; S[p][o] = sop;
            S[o][q] = soq; S[q][o] = soq;
            for (int k = 0; k < 3; k++) {
                const float vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
}

/* null vector of the 8 x 9 system whose transpose is M (9 x 8): Householder
 * QR of M, f = Q e9; the GPU's fgeom::null_vector_9x8 step for step */
static void null_vector_9x8_f32(float M[9][8], float f[9])
{
    /* Householder QR of M; reflector k stored in vk[k][k..8], beta[k] */
    float vk[8][9];
    float beta[8];
    for (int k = 0; k < 8; k++) {
        float nrm2 = 0.f;
        for (int r = k; r < 9; r++) nrm2 = nrm2 + M[r][k] * M[r][k];
        const float nrm = sqrtf(nrm2);
        for (int r = 0; r < 9; r++) vk[k][r] = 0.f;
        if (nrm == 0.f) { beta[k] = 0.f; continue; }
        const float alpha = (M[k][k] > 0.f) ? -nrm : nrm;
        vk[k][k] = M[k][k] - alpha;
        for (int r = k + 1; r < 9; r++) vk[k][r] = M[r][k];
        float vtv = 0.f;
        for (int r = k; r < 9; r++) vtv = vtv + vk[k][r] * vk[k][r];
        if (vtv == 0.f) { beta[k] = 0.f; continue; }
        beta[k] = 2.f / vtv;
        for (int j = k + 1; j < 8; j++) {
            float w = 0.f;
            for (int r = k; r < 9; r++) w = w + vk[k][r] * M[r][j];
            const float bw = beta[k] * w;
            for (int r = k; r < 9; r++) M[r][j] = M[r][j] - bw * vk[k][r];
        }
    }
    for (int r = 0; r < 9; r++) f[r] = r == 8 ? 1.f : 0.f;
    for (int k = 7; k >= 0; k--) {
        if (beta[k] == 0.f) continue;
        float w = 0.f;
        for (int r = k; r < 9; r++) w = w + vk[k][r] * f[r];
        const float bw = beta[k] * w;
        for (int r = k; r < 9; r++) f[r] = f[r] - bw * vk[k][r];
    }
}

int or_f8_fit_f32(const float* pts8, float F[9])
{
    /* Hartley normalisation */
    float c[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 4; k++) c[k] = c[k] + pts8[i * 4 + k];
    for (int k = 0; k < 4; k++) c[k] = c[k] / 8.f;
    float n0 = 0.f, n1 = 0.f;
    float d[8][4];
    for (int i = 0; i < 8; i++) {
        for (int k = 0; k < 4; k++) d[i][k] = pts8[i * 4 + k] - c[k];
        n0 = n0 + sqrtf(d[i][0] * d[i][0] + d[i][1] * d[i][1]);
        n1 = n1 + sqrtf(d[i][2] * d[i][2] + d[i][3] * d[i][3]);
    }
    n0 = n0 / 8.f;
    n1 = n1 / 8.f;
    if (!(n0 > 0.f) || !(n1 > 0.f)) return 0;
    const float sqrt2 = 1.41421354f;
    const float s0 = sqrt2 / n0, s1 = sqrt2 / n1;
    /* M = P^T, 9 x 8: column i is the design row of point i */
    float M[9][8];
    for (int i = 0; i < 8; i++) {
        const float x0 = d[i][0] * s0, y0 = d[i][1] * s0;
        const float x1 = d[i][2] * s1, y1 = d[i][3] * s1;
        M[0][i] = x1 * x0; M[1][i] = y1 * x0; M[2][i] = x0;
        M[3][i] = x1 * y0; M[4][i] = y1 * y0; M[5][i] = y0;
        M[6][i] = x1;      M[7][i] = y1;      M[8][i] = 1.f;
    }
    float f[9];
    null_vector_9x8_f32(M, f);
    /* F^(r,c) = f[3c + r] (column-major) */
    float Fh[3][3];
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) Fh[r][cc] = f[3 * cc + r];
    /* rank 2: smallest eigenvector of Fh^T Fh */
    float S[3][3], V[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i][j] = (Fh[0][i] * Fh[0][j] + Fh[1][i] * Fh[1][j]) + Fh[2][i] * Fh[2][j];
    jacobi3_f32(S, V);
    int m = 0;
    if (S[1][1] < S[m][m]) m = 1;
    if (S[2][2] < S[m][m]) m = 2;
    const float v3[3] = {V[0][m], V[1][m], V[2][m]};
    for (int r = 0; r < 3; r++) {
        const float fv = (Fh[r][0] * v3[0] + Fh[r][1] * v3[1]) + Fh[r][2] * v3[2];
        for (int cc = 0; cc < 3; cc++) Fh[r][cc] = Fh[r][cc] - fv * v3[cc];
    }
    /* denormalise: F = T1^T Fh T0 */
    const float a0 = -(s0 * c[0]), b0 = -(s0 * c[1]);
    const float a1 = -(s1 * c[2]), b1 = -(s1 * c[3]);
    float G[3][3];
    for (int r = 0; r < 3; r++) {
        G[r][0] = Fh[r][0] * s0;
        G[r][1] = Fh[r][1] * s0;
        G[r][2] = (Fh[r][0] * a0 + Fh[r][1] * b0) + Fh[r][2];
    }
    float R[3][3];
    for (int cc = 0; cc < 3; cc++) {
        R[0][cc] = s1 * G[0][cc];
        R[1][cc] = s1 * G[1][cc];
        R[2][cc] = (a1 * G[0][cc] + b1 * G[1][cc]) + G[2][cc];
    }
    float fro = 0.f;
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) fro = fro + R[r][cc] * R[r][cc];
    fro = sqrtf(fro);
    if (!(fro > 0.f) || !isfinite(fro)) return 0;
    const float inv = 1.f / fro;
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) F[r * 3 + cc] = R[r][cc] * inv;
    return 1;
}

int or_f_inlier_f32(const float F[9], const float* p, float t, int mode)
{
    const float x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3];
    const float t2 = t * t;
    if (mode == 0) {
        /* epipole0 = x1^T F, epipole1 = F x0 (geometry.cpp:66-75), no contraction */
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
    }
    /* Sampson (own test): the GPU's fused multiply-adds, one rounding each */
    const float l0a = fmaf(x1, F[0], fmaf(y1, F[3], F[6]));
    const float l0b = fmaf(x1, F[1], fmaf(y1, F[4], F[7]));
    const float l0c = fmaf(x1, F[2], fmaf(y1, F[5], F[8]));
    const float l1a = fmaf(x0, F[0], fmaf(y0, F[1], F[2]));
    const float l1b = fmaf(x0, F[3], fmaf(y0, F[4], F[5]));
    const float e = fmaf(l0a, x0, fmaf(l0b, y0, l0c));
    const float n = fmaf(l0a, l0a, fmaf(l0b, l0b, fmaf(l1a, l1a, l1b * l1b)));
    return e * e < t2 * n;
}

/* ---------------- FP64 oracle fit ---------------- */

static void jacobi_sym_f64(int n, double* A, double* V)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) off += A[i * n + j] * A[i * n + j];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0) t = -t;
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
}

int or_f8_fit_f64(const double* pts8, double F[9])
{
    double c[4] = {0, 0, 0, 0};
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 4; k++) c[k] += pts8[i * 4 + k] / 8.0;
    double n0 = 0, n1 = 0, d[8][4];
    for (int i = 0; i < 8; i++) {
        for (int k = 0; k < 4; k++) d[i][k] = pts8[i * 4 + k] - c[k];
        n0 += sqrt(d[i][0] * d[i][0] + d[i][1] * d[i][1]) / 8.0;
        n1 += sqrt(d[i][2] * d[i][2] + d[i][3] * d[i][3]) / 8.0;
    }
    if (!(n0 > 0) || !(n1 > 0)) return 0;
    const double s0 = sqrt(2.0) / n0, s1 = sqrt(2.0) / n1;
    double P[8][9];
    for (int i = 0; i < 8; i++) {
        const double x0 = d[i][0] * s0, y0 = d[i][1] * s0, x1 = d[i][2] * s1, y1 = d[i][3] * s1;
        const double row[9] = {x1 * x0, y1 * x0, x0, x1 * y0, y1 * y0, y0, x1, y1, 1};
        memcpy(P[i], row, sizeof(row));
    }
    double A[81], V[81];
    for (int a = 0; a < 9; a++)
        for (int b = 0; b < 9; b++) {
            double s = 0;
            for (int i = 0; i < 8; i++) s += P[i][a] * P[i][b];
            A[a * 9 + b] = s;
        }
    jacobi_sym_f64(9, A, V);
    int m = 0;
    for (int i = 1; i < 9; i++)
        if (A[i * 9 + i] < A[m * 9 + m]) m = i;
    double f[9];
    for (int i = 0; i < 9; i++) f[i] = V[i * 9 + m];
    double Fh[3][3];
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) Fh[r][cc] = f[3 * cc + r];
    double S[9], W[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i * 3 + j] = Fh[0][i] * Fh[0][j] + Fh[1][i] * Fh[1][j] + Fh[2][i] * Fh[2][j];
    jacobi_sym_f64(3, S, W);
    int q = 0;
    if (S[4] < S[q * 4]) q = 1;
    if (S[8] < S[q * 4]) q = 2;
    const double v3[3] = {W[0 * 3 + q], W[1 * 3 + q], W[2 * 3 + q]};
    for (int r = 0; r < 3; r++) {
        const double fv = Fh[r][0] * v3[0] + Fh[r][1] * v3[1] + Fh[r][2] * v3[2];
        for (int cc = 0; cc < 3; cc++) Fh[r][cc] -= fv * v3[cc];
    }
    const double T0[9] = {s0, 0, -s0 * c[0], 0, s0, -s0 * c[1], 0, 0, 1};
    const double T1[9] = {s1, 0, -s1 * c[2], 0, s1, -s1 * c[3], 0, 0, 1};
    double G[9], R[9];
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += Fh[r][k] * T0[k * 3 + cc];
            G[r * 3 + cc] = s;
        }
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += T1[k * 3 + r] * G[k * 3 + cc];
            R[r * 3 + cc] = s;
        }
    double fro = 0;
    for (int i = 0; i < 9; i++) fro += R[i] * R[i];
    fro = sqrt(fro);
    if (!(fro > 0)) return 0;
    for (int i = 0; i < 9; i++) F[i] = R[i] / fro;
    return 1;
}

int or_f_inlier_f64(const double F[9], const double* p, double t, int mode)
{
    const double x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3];
    const double l0a = x1 * F[0] + y1 * F[3] + F[6];
    const double l0b = x1 * F[1] + y1 * F[4] + F[7];
    const double l0c = x1 * F[2] + y1 * F[5] + F[8];
    const double l1a = x0 * F[0] + y0 * F[1] + F[2];
    const double l1b = x0 * F[3] + y0 * F[4] + F[5];
    const double e = l0a * x0 + l0b * y0 + l0c;
    const double n0 = l0a * l0a + l0b * l0b, n1 = l1a * l1a + l1b * l1b;
    if (mode == 0) return (e * e < t * t * n0) && (e * e < t * t * n1);
    return e * e < t * t * (n0 + n1);
}

/* ---------------- hypothesis scoring / drivers ---------------- */

static int sample_sane(const float* pts, const uint32_t* idx, float min_sep_sq0, float min_sep_sq1)
{
    if (min_sep_sq0 <= 0.f && min_sep_sq1 <= 0.f) return 1;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < i; j++) {
            const float* a = pts + (size_t)idx[i] * 4;
            const float* b = pts + (size_t)idx[j] * 4;
            const float dx0 = a[0] - b[0], dy0 = a[1] - b[1], dx1 = a[2] - b[2], dy1 = a[3] - b[3];
            if (dx0 * dx0 + dy0 * dy0 < min_sep_sq0 || dx1 * dx1 + dy1 * dy1 < min_sep_sq1) return 0;
        }
    return 1;
}

static uint32_t score_one(const float* pts, uint32_t n, const uint32_t* idx, float t, int mode,
                          float min_sep_sq0, float min_sep_sq1, float* model)
{
    /* fewer than 8 points or an index out of range: the hypothesis is rejected */
    int in_range = n >= 8;
    for (int i = 0; i < 8; i++) in_range &= idx[i] < n;
    if (!in_range || !sample_sane(pts, idx, min_sep_sq0, min_sep_sq1)) {
        if (model) for (int q = 0; q < 9; q++) model[q] = 0.f;
        return 0;
    }
    float p8[32], F[9];
    for (int i = 0; i < 8; i++) memcpy(p8 + i * 4, pts + (size_t)idx[i] * 4, 16);
    if (!or_f8_fit_f32(p8, F)) {
        if (model) for (int q = 0; q < 9; q++) model[q] = 0.f;
        return 0;
    }
    if (model) memcpy(model, F, sizeof(F));
    uint32_t cnt = 0;
    for (uint32_t k = 0; k < n; k++) cnt += (uint32_t)or_f_inlier_f32(F, pts + (size_t)k * 4, t, mode);
    return cnt;
}

void or_f8_score(const float* pts, uint32_t n, const uint32_t* idx, uint32_t H, float t, int mode,
                 float min_sep_sq0, float min_sep_sq1, uint32_t* counts, float* models, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t h = 0; h < (int64_t)H; h++)
        counts[h] = score_one(pts, n, idx + (size_t)h * 8, t, mode, min_sep_sq0, min_sep_sq1,
                              models ? models + (size_t)h * 9 : NULL);
}

uint32_t or_ransac_required_iters(uint32_t best, uint32_t n, uint32_t k, double conf, uint32_t max_iter)
{
    /* ransac.hpp:155-157 (confidence is a float there) */
    const double num = log2(1.0 - (double)(float)conf);
    const double den = log2(1.f - pow((double)best / n, (double)k));
    const double r = ceil(num / den);
    const uint64_t req = (uint64_t)r;
    return (uint32_t)(req < max_iter ? req : max_iter);
}

typedef uint32_t (*score_fn)(const uint32_t* idx, void* ctx);

/* ransac<Func, 8> of ransac.hpp:115-163 with its TableSampler (ransac.hpp:52-82) */
static uint32_t ransac_loop(uint32_t n, float confidence, uint32_t max_iter, uint64_t seed,
                            score_fn fn, void* ctx, uint32_t best_idx[8], uint32_t* draws_out)
{
    or_mt64 rng;
    or_mt64_seed(&rng, seed);
    uint8_t* mark = (uint8_t*)calloc(n ? n : 1, 1);
    uint32_t best = 0;
    for (uint32_t i = 0; i < 8; i++) best_idx[i] = i;
    uint32_t required = max_iter, iters = 0, insane = 0, draws = 0;
    const uint32_t insane_cap = max_iter < 128 ? max_iter : 128;
    do {
        uint32_t idx[8];
        for (int i = 0; i < 8; i++) {
            uint32_t v;
            do { v = (uint32_t)(or_mt64_next(&rng) % n); } while (mark[v]);
            idx[i] = v;
            mark[v] = 1;
        }
        for (int i = 0; i < 8; i++) mark[idx[i]] = 0;
        draws++;
        const uint32_t r = fn(idx, ctx);
        if (r == 0) {
            insane++;
            if (insane > insane_cap) break;
            continue;
        }
        insane = 0;
        if (r > best) {
            best = r;
            memcpy(best_idx, idx, sizeof(idx));
            required = or_ransac_required_iters(best, n, 8, confidence, max_iter);
        }
        iters++;
    } while (iters < required);
    free(mark);
    if (draws_out) *draws_out = draws;
    return best;
}

typedef struct { const float* pts; uint32_t n; float t; int mode; float s0, s1; } f8_ctx;

static uint32_t f8_score_fn(const uint32_t* idx, void* vctx)
{
    const f8_ctx* c = (const f8_ctx*)vctx;
    return score_one(c->pts, c->n, idx, c->t, c->mode, c->s0, c->s1, NULL);
}

uint32_t or_f8_ransac_adaptive(const float* pts, uint32_t n, float t, int mode, float confidence,
                               uint32_t max_iter, uint64_t seed, float min_sep_sq0, float min_sep_sq1,
                               uint32_t best_idx[8], uint32_t* iters_out)
{
    f8_ctx c = {pts, n, t, mode, min_sep_sq0, min_sep_sq1};
    return ransac_loop(n, confidence, max_iter, seed, f8_score_fn, &c, best_idx, iters_out);
}

typedef struct { uint32_t modulus; } hash_ctx;

static uint32_t hash_score_fn(const uint32_t* idx, void* vctx)
{
    const hash_ctx* c = (const hash_ctx*)vctx;
    uint64_t s = 0;
    for (int i = 0; i < 8; i++) s = s * 1000003ULL + idx[i];
    return (uint32_t)((s * 2654435761ULL) % c->modulus);
}

uint32_t or_ransac_loop_hash(uint32_t n, float confidence, uint32_t max_iter, uint64_t seed,
                             uint32_t modulus, uint32_t best_idx[8], uint32_t* draws_out)
{
    hash_ctx c = {modulus};
    return ransac_loop(n, confidence, max_iter, seed, hash_score_fn, &c, best_idx, draws_out);
}


/* ---------------- 4-point homography hypotheses ---------------- */

/* findHomography (legacy/geometry.hpp:139-209) on 4 correspondences, the GPU's
 * fgeom::fit_h4 step for step: Hartley normalisation as in or_f8_fit_f32, DLT
 * rows [-x,-y,-1,0,0,0,x x',y x',x'] / [0,0,0,-x,-y,-1,x y',y y',y'], the same
 * Householder null vector, H = T1^-1 H^ T0 (row-major), unit Frobenius norm. */
int or_h4_fit_f32(const float* pts4, float H[9])
{
    float c[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) c[k] = c[k] + pts4[i * 4 + k];
    for (int k = 0; k < 4; k++) c[k] = c[k] / 4.f;
    float n0 = 0.f, n1 = 0.f;
    float d[4][4];
    for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 4; k++) d[i][k] = pts4[i * 4 + k] - c[k];
        n0 = n0 + sqrtf(d[i][0] * d[i][0] + d[i][1] * d[i][1]);
        n1 = n1 + sqrtf(d[i][2] * d[i][2] + d[i][3] * d[i][3]);
    }
    n0 = n0 / 4.f;
    n1 = n1 / 4.f;
    if (!(n0 > 0.f) || !(n1 > 0.f)) return 0;
    const float sqrt2 = 1.41421354f;
    const float s0 = sqrt2 / n0, s1 = sqrt2 / n1;
    float M[9][8];
    for (int i = 0; i < 4; i++) {
        const float x = d[i][0] * s0, y = d[i][1] * s0;
        const float xp = d[i][2] * s1, yp = d[i][3] * s1;
        const int a = 2 * i, b = 2 * i + 1;
        M[0][a] = -x;  M[1][a] = -y;  M[2][a] = -1.f;
        M[3][a] = 0.f; M[4][a] = 0.f; M[5][a] = 0.f;
        M[6][a] = x * xp; M[7][a] = y * xp; M[8][a] = xp;
        M[0][b] = 0.f; M[1][b] = 0.f; M[2][b] = 0.f;
        M[3][b] = -x;  M[4][b] = -y;  M[5][b] = -1.f;
        M[6][b] = x * yp; M[7][b] = y * yp; M[8][b] = yp;
    }
    float h[9];
    null_vector_9x8_f32(M, h);
    const float a0 = -(s0 * c[0]), b0 = -(s0 * c[1]);
    float G[3][3], R[3][3];
    for (int r = 0; r < 3; r++) {
        G[r][0] = h[3 * r + 0] * s0;
        G[r][1] = h[3 * r + 1] * s0;
        G[r][2] = (h[3 * r + 0] * a0 + h[3 * r + 1] * b0) + h[3 * r + 2];
    }
    const float inv1 = 1.f / s1;
    for (int cc = 0; cc < 3; cc++) {
        R[0][cc] = G[0][cc] * inv1 + c[2] * G[2][cc];
        R[1][cc] = G[1][cc] * inv1 + c[3] * G[2][cc];
        R[2][cc] = G[2][cc];
    }
    float fro = 0.f;
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) fro = fro + R[r][cc] * R[r][cc];
    fro = sqrtf(fro);
    if (!(fro > 0.f) || !isfinite(fro)) return 0;
    const float inv = 1.f / fro;
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) H[r * 3 + cc] = R[r][cc] * inv;
    return 1;
}

/* findHomographyRansacWithInsufficientTiePts<0>'s per-hypothesis body
 * (ImagePair.cpp:365-378) over given [H][4] samples: sanity gate, 4-point fit,
 * checkProjectiveTransformation count */
void or_h4_score(const float* pts, uint32_t n, const uint32_t* idx, uint32_t H, float t, float min_sep_sq0,
                 float min_sep_sq1, uint32_t* counts, float* models, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t h = 0; h < (int64_t)H; h++) {
        const uint32_t* id = idx + (size_t)h * 4;
        int ok = n >= 4;
        for (int i = 0; i < 4; i++) ok &= id[i] < n;
        if (ok && (min_sep_sq0 > 0.f || min_sep_sq1 > 0.f))
            for (int i = 0; i < 4 && ok; i++)
                for (int j = 0; j < i; j++) {
                    const float* a = pts + (size_t)id[i] * 4;
                    const float* b = pts + (size_t)id[j] * 4;
                    const float dx0 = a[0] - b[0], dy0 = a[1] - b[1], dx1 = a[2] - b[2], dy1 = a[3] - b[3];
                    if (dx0 * dx0 + dy0 * dy0 < min_sep_sq0 || dx1 * dx1 + dy1 * dy1 < min_sep_sq1) ok = 0;
                }
        float Hm[9] = {0};
        if (ok) {
            float p4[16];
            for (int i = 0; i < 4; i++) memcpy(p4 + 4 * i, pts + (size_t)id[i] * 4, 16);
            ok = or_h4_fit_f32(p4, Hm);
        }
        uint32_t cnt = 0;
        if (ok)
            for (uint32_t k = 0; k < n; k++) cnt += (uint32_t)or_h_inlier_f32(Hm, pts + (size_t)k * 4, t);
        counts[h] = ok ? cnt : 0u;
        if (models)
            for (int q = 0; q < 9; q++) models[(size_t)h * 9 + q] = ok ? Hm[q] : 0.f;
    }
}
