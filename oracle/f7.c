/*
 * f7.c - 7-point fundamental-matrix hypotheses restated on the CPU.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * Reference: findEpipolarityWith7Points (legacy/geometry.cpp:221-307) as the
 * model of findEpipolarityDegensac (degensac.cpp:113-150): Hartley
 * normalisation, the 7x9 design rows of the 8-point fit, the two null vectors
 * F1, F2, real roots of det(a F1 + (1 - a) F2) = 0 (the reference solves the
 * cubic by companion-matrix eigenvalues), F = T1^T F^ T0.
 *
 *  - or_f7_fit_f32 repeats the GPU solver step for step (FP32 Householder QR
 *    of the transposed system for the null space, exact FP64 cubic
 *    coefficients from four determinants, FP64 bisection roots between the
 *    critical points, unit Frobenius norm), so GPU and CPU models and counts
 *    are identical. Build with -ffp-contract=off.
 *  - or_f7_fit_f64: the same algorithm in FP64 throughout (accuracy oracle).
 * Pinned in tests by noise-free geometry (one solution equals the true F up
 * to sign), the seven epipolar constraints and det F = 0 of every solution,
 * and the number of real roots against numpy's companion-matrix roots of the
 * same cubic. The reference's Eigen eigenvalue route is not reproduced bit for
 * bit (Eigen is absent, so its own FP32 bits cannot be produced here).
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double det3d(const double G[3][3])
{
    return G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
           G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
}

static double cubic_at(const double a[4], double t) { return ((a[3] * t + a[2]) * t + a[1]) * t + a[0]; }

static double bisect(const double a[4], double lo, double hi)
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

static int cubic_roots(const double a[4], double r[3])
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

/* cubic coefficients of det(F2 + t D) from four exact evaluations; 0 if degenerate */
static int cubic_of(const double F2[3][3], const double D[3][3], double a[4])
{
    const double ts[4] = {0.0, 1.0, -1.0, 2.0};
    double d[4];
    for (int q = 0; q < 4; q++) {
        double G[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) G[r][c] = F2[r][c] + ts[q] * D[r][c];
        d[q] = det3d(G);
    }
    a[0] = d[0];
    a[2] = 0.5 * (d[1] + d[2]) - d[0];
    const double sOdd = 0.5 * (d[1] - d[2]);
    a[3] = ((d[3] - d[0]) - 4.0 * a[2] - 2.0 * sOdd) / 6.0;
    a[1] = sOdd - a[3];
    double scale = fabs(a[0]);
    for (int q = 1; q < 4; q++) scale = fabs(a[q]) > scale ? fabs(a[q]) : scale;
    return (scale > 0.0) && (fabs(a[3]) > 1e-12 * scale);
}

int or_f7_fit_f32(const float* pts7, float* Fs)
{
    float c[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 7; i++) {
        c[0] = c[0] + pts7[4 * i];
        c[1] = c[1] + pts7[4 * i + 1];
        c[2] = c[2] + pts7[4 * i + 2];
        c[3] = c[3] + pts7[4 * i + 3];
    }
    for (int k = 0; k < 4; k++) c[k] = c[k] / 7.f;
    float n0 = 0.f, n1 = 0.f;
    for (int i = 0; i < 7; i++) {
        const float d0 = pts7[4 * i] - c[0], d1 = pts7[4 * i + 1] - c[1];
        const float d2 = pts7[4 * i + 2] - c[2], d3 = pts7[4 * i + 3] - c[3];
        n0 = n0 + sqrtf(d0 * d0 + d1 * d1);
        n1 = n1 + sqrtf(d2 * d2 + d3 * d3);
    }
    n0 = n0 / 7.f;
    n1 = n1 / 7.f;
    if (!(n0 > 0.f) || !(n1 > 0.f)) return 0;
    const float sqrt2 = 1.41421354f;
    const float s0 = sqrt2 / n0, s1 = sqrt2 / n1;
    float M[9][7];
    for (int i = 0; i < 7; i++) {
        const float x0 = (pts7[4 * i] - c[0]) * s0, y0 = (pts7[4 * i + 1] - c[1]) * s0;
        const float x1 = (pts7[4 * i + 2] - c[2]) * s1, y1 = (pts7[4 * i + 3] - c[3]) * s1;
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
    for (int k = 0; k < 7; k++) {
        float nrm2 = 0.f;
        for (int r = k; r < 9; r++) nrm2 = nrm2 + M[r][k] * M[r][k];
        const float nrm = sqrtf(nrm2);
        vkk[k] = 0.f;
        beta[k] = 0.f;
        if (nrm == 0.f) {
            for (int r = k + 1; r < 9; r++) M[r][k] = 0.f;
            continue;
        }
        const float alpha = (M[k][k] > 0.f) ? -nrm : nrm;
        vkk[k] = M[k][k] - alpha;
        float vtv = vkk[k] * vkk[k];
        for (int r = k + 1; r < 9; r++) vtv = vtv + M[r][k] * M[r][k];
        if (vtv == 0.f) continue;
        beta[k] = 2.f / vtv;
        for (int j = k + 1; j < 7; j++) {
            float w = vkk[k] * M[k][j];
            for (int r = k + 1; r < 9; r++) w = w + M[r][k] * M[r][j];
            const float bw = beta[k] * w;
            M[k][j] = M[k][j] - bw * vkk[k];
            for (int r = k + 1; r < 9; r++) M[r][j] = M[r][j] - bw * M[r][k];
        }
    }
    float f[2][9];
    for (int q = 0; q < 2; q++) {
        for (int r = 0; r < 9; r++) f[q][r] = (r == 8 - q) ? 1.f : 0.f;
        for (int k = 6; k >= 0; k--) {
            if (beta[k] == 0.f) continue;
            float w = vkk[k] * f[q][k];
            for (int r = k + 1; r < 9; r++) w = w + M[r][k] * f[q][r];
            const float bw = beta[k] * w;
            f[q][k] = f[q][k] - bw * vkk[k];
            for (int r = k + 1; r < 9; r++) f[q][r] = f[q][r] - bw * M[r][k];
        }
    }
    double F2[3][3], D[3][3];
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) {
            F2[r][cc] = (double)f[1][3 * cc + r];
            D[r][cc] = (double)f[0][3 * cc + r] - F2[r][cc];
        }
    double a[4];
    if (!cubic_of(F2, D, a)) return 0;
    double roots[3];
    const int nr = cubic_roots(a, roots);
    const float a0 = -(s0 * c[0]), b0 = -(s0 * c[1]);
    const float a1 = -(s1 * c[2]), b1 = -(s1 * c[3]);
    int out = 0;
    for (int k = 0; k < nr; k++) {
        float Fh[3][3], G[3][3], R[3][3];
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) Fh[r][cc] = (float)(F2[r][cc] + roots[k] * D[r][cc]);
        for (int r = 0; r < 3; r++) {
            G[r][0] = Fh[r][0] * s0;
            G[r][1] = Fh[r][1] * s0;
            G[r][2] = (Fh[r][0] * a0 + Fh[r][1] * b0) + Fh[r][2];
        }
        for (int cc = 0; cc < 3; cc++) {
            R[0][cc] = s1 * G[0][cc];
            R[1][cc] = s1 * G[1][cc];
            R[2][cc] = (a1 * G[0][cc] + b1 * G[1][cc]) + G[2][cc];
        }
        float fro = 0.f;
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) fro = fro + R[r][cc] * R[r][cc];
        fro = sqrtf(fro);
        if (!(fro > 0.f) || !isfinite(fro)) continue;
        const float invf = 1.f / fro;
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) Fs[out * 9 + r * 3 + cc] = R[r][cc] * invf;
        out++;
    }
    return out;
}

/* FP64: centroid/scale, null space from the eigenvectors of P^T P (Jacobi to
 * convergence), the same cubic and roots, denormalised, unit Frobenius */
static void jacobi_sym(int n, double* A, double* V)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[i * n + j] = i == j;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                const double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                double t = 1 / (fabs(th) + sqrt(th * th + 1));
                if (th < 0) t = -t;
                const double cs = 1 / sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < n; k++) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = cs * akp - sn * akq;
                    A[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = cs * apk - sn * aqk;
                    A[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = cs * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
}

int or_f7_fit_f64(const double* pts7, double* Fs)
{
    double c[4] = {0, 0, 0, 0};
    for (int i = 0; i < 7; i++)
        for (int k = 0; k < 4; k++) c[k] += pts7[4 * i + k];
    for (int k = 0; k < 4; k++) c[k] /= 7;
    double n0 = 0, n1 = 0;
    for (int i = 0; i < 7; i++) {
        n0 += hypot(pts7[4 * i] - c[0], pts7[4 * i + 1] - c[1]);
        n1 += hypot(pts7[4 * i + 2] - c[2], pts7[4 * i + 3] - c[3]);
    }
    n0 /= 7;
    n1 /= 7;
    if (!(n0 > 0) || !(n1 > 0)) return 0;
    const double s0 = sqrt(2.0) / n0, s1 = sqrt(2.0) / n1;
    double P[7][9];
    for (int i = 0; i < 7; i++) {
        const double x0 = (pts7[4 * i] - c[0]) * s0, y0 = (pts7[4 * i + 1] - c[1]) * s0;
        const double x1 = (pts7[4 * i + 2] - c[2]) * s1, y1 = (pts7[4 * i + 3] - c[3]) * s1;
        const double row[9] = {x1 * x0, y1 * x0, x0, x1 * y0, y1 * y0, y0, x1, y1, 1};
        memcpy(P[i], row, sizeof(row));
    }
    double A[81], V[81];
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 9; j++) {
            double s = 0;
            for (int k = 0; k < 7; k++) s += P[k][i] * P[k][j];
            A[i * 9 + j] = s;
        }
    jacobi_sym(9, A, V);
    /* the two smallest eigenvalues span the null space */
    int i1 = 0, i2 = -1;
    for (int i = 1; i < 9; i++)
        if (A[i * 9 + i] < A[i1 * 9 + i1]) i1 = i;
    for (int i = 0; i < 9; i++)
        if (i != i1 && (i2 < 0 || A[i * 9 + i] < A[i2 * 9 + i2])) i2 = i;
    double F1[3][3], F2[3][3], D[3][3];
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) {
            F1[r][cc] = V[(3 * cc + r) * 9 + i1];
            F2[r][cc] = V[(3 * cc + r) * 9 + i2];
        }
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) D[r][cc] = F1[r][cc] - F2[r][cc];
    double a[4];
    if (!cubic_of(F2, D, a)) return 0;
    double roots[3];
    const int nr = cubic_roots(a, roots);
    int out = 0;
    for (int k = 0; k < nr; k++) {
        double Fh[3][3], R[3][3];
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) Fh[r][cc] = F2[r][cc] + roots[k] * D[r][cc];
        /* R = T1^T Fh T0, T = [s 0 -s cx; 0 s -s cy; 0 0 1] */
        const double T0[3][3] = {{s0, 0, -s0 * c[0]}, {0, s0, -s0 * c[1]}, {0, 0, 1}};
        const double T1[3][3] = {{s1, 0, -s1 * c[2]}, {0, s1, -s1 * c[3]}, {0, 0, 1}};
        double G[3][3];
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                double s = 0;
                for (int k2 = 0; k2 < 3; k2++) s += Fh[r][k2] * T0[k2][cc];
                G[r][cc] = s;
            }
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) {
                double s = 0;
                for (int k2 = 0; k2 < 3; k2++) s += T1[k2][r] * G[k2][cc];
                R[r][cc] = s;
            }
        double fro = 0;
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) fro += R[r][cc] * R[r][cc];
        fro = sqrt(fro);
        if (!(fro > 0)) continue;
        for (int r = 0; r < 3; r++)
            for (int cc = 0; cc < 3; cc++) Fs[out * 9 + r * 3 + cc] = R[r][cc] / fro;
        out++;
    }
    return out;
}

static int sane7(const float* pts, const uint32_t* idx, float s0, float s1)
{
    if (s0 <= 0.f && s1 <= 0.f) return 1;
    for (int i = 0; i < 7; i++)
        for (int j = 0; j < i; j++) {
            const float* a = pts + (size_t)idx[i] * 4;
            const float* b = pts + (size_t)idx[j] * 4;
            const float dx0 = a[0] - b[0], dy0 = a[1] - b[1], dx1 = a[2] - b[2], dy1 = a[3] - b[3];
            if (dx0 * dx0 + dy0 * dy0 < s0 || dx1 * dx1 + dy1 * dy1 < s1) return 0;
        }
    return 1;
}

void or_f7_score(const float* pts, uint32_t n, const uint32_t* idx, uint32_t H, float t, int mode, float min_sep_sq0,
                 float min_sep_sq1, uint32_t* counts, float* models, int nthreads)
{
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t h = 0; h < (int64_t)H; h++) {
        const uint32_t* id = idx + (size_t)h * 7;
        int ok = n >= 7;
        for (int i = 0; i < 7; i++) ok &= id[i] < n;
        float Fs[27];
        int nm = 0;
        if (ok && sane7(pts, id, min_sep_sq0, min_sep_sq1)) {
            float p7[28];
            for (int i = 0; i < 7; i++) memcpy(p7 + 4 * i, pts + (size_t)id[i] * 4, 16);
            nm = or_f7_fit_f32(p7, Fs);
        }
        uint32_t best = 0;
        int bm = 0;
        for (int m = 0; m < nm; m++) {
            uint32_t cnt = 0;
            for (uint32_t k = 0; k < n; k++) cnt += (uint32_t)or_f_inlier_f32(Fs + 9 * m, pts + (size_t)k * 4, t, mode);
            if (m == 0 || cnt > best) {
                best = cnt;
                bm = m;
            }
        }
        counts[h] = nm ? best : 0u;
        if (models)
            for (int q = 0; q < 9; q++) models[(size_t)h * 9 + q] = nm ? Fs[9 * bm + q] : 0.f;
    }
}
