/*
 * degensac.c - the DEGENSAC finisher restated on the CPU.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * Reference: findEpipolarityDegensac (degensac.cpp:102-240) with
 * checkEpipolarityAndOptimise / checkHomographyAndOptimise
 * (degensac.cpp:57-100), findHomographyFromEpipolarity
 * (legacy/geometry.cpp:311-327), findEpipolarityFromHomography
 * (legacy/geometry.cpp:330-339), checkProjectiveTransformation
 * (legacy/geometry.cpp:41-53), findEpipolarity / findHomography
 * (legacy/geometry.hpp:139-259).
 *
 * The GPU finisher (rsfm_degensac_refine) is repeated step for step so that
 * its models, counts and masks are identical: the candidate order, the
 * evenly-by-rank inlier subsample of at most 128, the FP64 least-squares
 * refits (normal matrix accumulated in point order, Cholesky of
 * N + 1e-10 tr(N) I, five inverse-iteration steps from the all-ones vector,
 * rank 2 by a cyclic 3x3 Jacobi), H from F (epipole = largest column cross
 * product, Cramer solve), F from H and two points, the FP32 homography test
 * and the plane-sample stream (splitmix64, ranks into the off-plane set).
 * Build with -ffp-contract=off.
 *
 * Pinned in tests by mathematics rather than by retyping: refits reproduce
 * noise-free F and H exactly up to scale and agree with an SVD null vector
 * (numpy) on noisy data; H from F maps its three points onto their partners
 * and satisfies F ~ [e']x H; F from H passes through both points, is rank 2
 * and makes H^T F antisymmetric; the homography test is checked against a
 * direct evaluation of the reference's formula; on a dominant-plane scene the
 * finisher recovers the true F where the plain 7-point best explains only
 * the plane. The reference's Bernoulli subsample (mt19937_64) and Eigen
 * JacobiSVD refits are not reproduced bit for bit, and the reference's own
 * trajectories cannot be compared: random_device seeding makes them
 * unrepeatable in the reference as well.
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DG_MAX_SEL 128
#define DG_MAX_PLANE 4096u
static const double kSqrt2 = 1.4142135623730951;
static const double kShift = 1e-10;

static const int kTri[5][3] = {{0, 1, 2}, {3, 4, 5}, {0, 1, 6}, {3, 4, 6}, {2, 5, 6}};
static const int kRem[5][4] = {{3, 4, 5, 6}, {0, 1, 2, 6}, {2, 3, 4, 5}, {0, 1, 2, 5}, {0, 1, 3, 4}};

static void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

static double det3(const double G[3][3])
{
    return G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
           G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
}

static void jacobi3d(double S[3][3], double V[3][3])
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

static int null_vector9(const double N[9][9], double v[9])
{
    double tr = 0.0;
    for (int a = 0; a < 9; a++) tr = tr + N[a][a];
    if (!(tr > 0.0) || !isfinite(tr)) return 0;
    const double sigma = kShift * tr;
    double L[9][9], rd[9]; /* rd: reciprocals of the diagonal, as the GPU */
    for (int j = 0; j < 9; j++) {
        double s = N[j][j] + sigma;
        for (int k = 0; k < j; k++) s = s - L[j][k] * L[j][k];
        if (!(s > 0.0)) return 0;
        L[j][j] = sqrt(s);
        rd[j] = 1.0 / L[j][j];
        for (int i = j + 1; i < 9; i++) {
            double t = N[i][j];
            for (int k = 0; k < j; k++) t = t - L[i][k] * L[j][k];
            L[i][j] = t * rd[j];
        }
    }
    for (int i = 0; i < 9; i++) v[i] = 1.0;
    for (int it = 0; it < 5; it++) {
        double y[9], z[9];
        for (int i = 0; i < 9; i++) {
            double t = v[i];
            for (int k = 0; k < i; k++) t = t - L[i][k] * y[k];
            y[i] = t * rd[i];
        }
        for (int i = 8; i >= 0; i--) {
            double t = y[i];
            for (int k = i + 1; k < 9; k++) t = t - L[k][i] * z[k];
            z[i] = t * rd[i];
        }
        double nrm2 = 0.0;
        for (int i = 0; i < 9; i++) nrm2 = nrm2 + z[i] * z[i];
        if (!(nrm2 > 0.0) || !isfinite(nrm2)) return 0;
        const double inv = 1.0 / sqrt(nrm2);
        for (int i = 0; i < 9; i++) v[i] = z[i] * inv;
    }
    return 1;
}

typedef struct {
    double c[4];
    double s0, s1;
} dg_norm;

static int to_unit_f32(const double R[3][3], float out[9])
{
    double fro = 0.0;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) fro = fro + R[r][c] * R[r][c];
    fro = sqrt(fro);
    if (!(fro > 0.0) || !isfinite(fro)) return 0;
    const double inv = 1.0 / fro;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) out[r * 3 + c] = (float)(R[r][c] * inv);
    return 1;
}

static int finish_f(const double f[9], const dg_norm* nm, float F[9])
{
    double Fh[3][3], S[3][3], V[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Fh[r][c] = f[3 * c + r];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i][j] = (Fh[0][i] * Fh[0][j] + Fh[1][i] * Fh[1][j]) + Fh[2][i] * Fh[2][j];
    jacobi3d(S, V);
    int m = 0;
    if (S[1][1] < S[0][0]) m = 1;
    if (S[2][2] < S[m][m]) m = 2;
    const double v3[3] = {V[0][m], V[1][m], V[2][m]};
    for (int r = 0; r < 3; r++) {
        const double fv = (Fh[r][0] * v3[0] + Fh[r][1] * v3[1]) + Fh[r][2] * v3[2];
        for (int c = 0; c < 3; c++) Fh[r][c] = Fh[r][c] - fv * v3[c];
    }
    const double a0 = -(nm->s0 * nm->c[0]), b0 = -(nm->s0 * nm->c[1]);
    const double a1 = -(nm->s1 * nm->c[2]), b1 = -(nm->s1 * nm->c[3]);
    double G[3][3], R[3][3];
    for (int r = 0; r < 3; r++) {
        G[r][0] = Fh[r][0] * nm->s0;
        G[r][1] = Fh[r][1] * nm->s0;
        G[r][2] = (Fh[r][0] * a0 + Fh[r][1] * b0) + Fh[r][2];
    }
    for (int c = 0; c < 3; c++) {
        R[0][c] = nm->s1 * G[0][c];
        R[1][c] = nm->s1 * G[1][c];
        R[2][c] = (a1 * G[0][c] + b1 * G[1][c]) + G[2][c];
    }
    return to_unit_f32(R, F);
}

static int finish_h(const double h[9], const dg_norm* nm, float H[9])
{
    const double a0 = -(nm->s0 * nm->c[0]), b0 = -(nm->s0 * nm->c[1]);
    double G[3][3], R[3][3];
    for (int r = 0; r < 3; r++) {
        G[r][0] = h[3 * r + 0] * nm->s0;
        G[r][1] = h[3 * r + 1] * nm->s0;
        G[r][2] = (h[3 * r + 0] * a0 + h[3 * r + 1] * b0) + h[3 * r + 2];
    }
    const double inv1 = 1.0 / nm->s1;
    for (int c = 0; c < 3; c++) {
        R[0][c] = G[0][c] * inv1 + nm->c[2] * G[2][c];
        R[1][c] = G[1][c] * inv1 + nm->c[3] * G[2][c];
        R[2][c] = G[2][c];
    }
    return to_unit_f32(R, H);
}

/* least-squares F (is_h = 0) or H (is_h = 1) on the m points sel[] of pts */
static int fit_ls(const float* pts, const uint32_t* sel, uint32_t m, int is_h, float out[9])
{
    double (*pt)[4] = malloc(sizeof(double[4]) * m);
    double (*row)[9] = malloc(sizeof(double[9]) * 2 * m);
    int ok = 0;
    for (uint32_t j = 0; j < m; j++)
        for (int k = 0; k < 4; k++) pt[j][k] = pts[(size_t)sel[j] * 4 + k];
    const double dm = (double)m;
    dg_norm nm;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t j = 0; j < m; j++)
        for (int k = 0; k < 4; k++) a[k] = a[k] + pt[j][k];
    for (int k = 0; k < 4; k++) nm.c[k] = a[k] / dm;
    double a0 = 0.0, a1 = 0.0;
    for (uint32_t j = 0; j < m; j++) {
        const double dx0 = pt[j][0] - nm.c[0], dy0 = pt[j][1] - nm.c[1];
        const double dx1 = pt[j][2] - nm.c[2], dy1 = pt[j][3] - nm.c[3];
        a0 = a0 + sqrt(dx0 * dx0 + dy0 * dy0);
        a1 = a1 + sqrt(dx1 * dx1 + dy1 * dy1);
    }
    const double n0 = a0 / dm, n1 = a1 / dm;
    if (!(n0 > 0.0) || !(n1 > 0.0)) goto done;
    nm.s0 = kSqrt2 / n0;
    nm.s1 = kSqrt2 / n1;
    for (uint32_t j = 0; j < m; j++) {
        const double x0 = (pt[j][0] - nm.c[0]) * nm.s0, y0 = (pt[j][1] - nm.c[1]) * nm.s0;
        const double x1 = (pt[j][2] - nm.c[2]) * nm.s1, y1 = (pt[j][3] - nm.c[3]) * nm.s1;
        if (is_h) {
            const double r1[9] = {-x0, -y0, -1.0, 0.0, 0.0, 0.0, x0 * x1, y0 * x1, x1};
            const double r2[9] = {0.0, 0.0, 0.0, -x0, -y0, -1.0, x0 * y1, y0 * y1, y1};
            memcpy(row[2 * j], r1, sizeof r1);
            memcpy(row[2 * j + 1], r2, sizeof r2);
        } else {
            const double r[9] = {x1 * x0, y1 * x0, x0, x1 * y0, y1 * y0, y0, x1, y1, 1.0};
            memcpy(row[j], r, sizeof r);
        }
    }
    {
        const uint32_t rows = is_h ? 2 * m : m;
        double N[9][9];
        for (int aa = 0; aa < 9; aa++)
            for (int b = aa; b < 9; b++) {
                double acc = 0.0;
                for (uint32_t r = 0; r < rows; r++) acc = acc + row[r][aa] * row[r][b];
                N[aa][b] = acc;
                N[b][aa] = acc;
            }
        double v[9];
        ok = null_vector9(N, v);
        if (ok) ok = is_h ? finish_h(v, &nm, out) : finish_f(v, &nm, out);
    }
done:
    free(pt);
    free(row);
    return ok;
}

int or_h_from_f(const float F[9], const float* pts3 /*3 x 4*/, float H[9])
{
    double Fd[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Fd[r][c] = (double)F[r * 3 + c];
    const double col[3][3] = {{Fd[0][0], Fd[1][0], Fd[2][0]}, {Fd[0][1], Fd[1][1], Fd[2][1]},
                              {Fd[0][2], Fd[1][2], Fd[2][2]}};
    double cand[3][3];
    cross3(col[0], col[1], cand[0]);
    cross3(col[0], col[2], cand[1]);
    cross3(col[1], col[2], cand[2]);
    int bi = 0;
    double bn = dot3(cand[0], cand[0]);
    for (int k = 1; k < 3; k++) {
        const double nk = dot3(cand[k], cand[k]);
        if (nk > bn) {
            bn = nk;
            bi = k;
        }
    }
    if (!(bn > 0.0) || !isfinite(bn)) return 0;
    const double inve = 1.0 / sqrt(bn);
    double e[3];
    for (int k = 0; k < 3; k++) e[k] = cand[bi][k] * inve;
    double A[3][3];
    for (int c = 0; c < 3; c++) {
        double ac[3];
        cross3(e, col[c], ac);
        A[0][c] = ac[0];
        A[1][c] = ac[1];
        A[2][c] = ac[2];
    }
    double b[3], M[3][3];
    for (int i = 0; i < 3; i++) {
        const float* p = pts3 + 4 * i;
        const double x[3] = {(double)p[0], (double)p[1], 1.0};
        const double xp[3] = {(double)p[2], (double)p[3], 1.0};
        double Ax[3], u[3], w[3];
        for (int r = 0; r < 3; r++) Ax[r] = (A[r][0] * x[0] + A[r][1] * x[1]) + A[r][2];
        cross3(xp, Ax, u);
        cross3(xp, e, w);
        const double ww = dot3(w, w);
        if (!(ww > 0.0)) return 0;
        b[i] = dot3(u, w) / ww;
        M[i][0] = x[0];
        M[i][1] = x[1];
        M[i][2] = 1.0;
    }
    const double dM = det3(M);
    if (!(dM != 0.0) || !isfinite(dM)) return 0;
    double v[3];
    for (int k = 0; k < 3; k++) {
        double Mk[3][3];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) Mk[r][c] = (c == k) ? b[r] : M[r][c];
        v[k] = det3(Mk) / dM;
    }
    double R[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) R[r][c] = A[r][c] - e[r] * v[c];
    return to_unit_f32(R, H);
}

int or_f_from_h(const float H[9], const float* pa, const float* pb, float F[9])
{
    double Hd[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) Hd[r][c] = (double)H[r * 3 + c];
    double l[2][3];
    for (int i = 0; i < 2; i++) {
        const float* q = i == 0 ? pa : pb;
        const double x0 = q[0], y0 = q[1];
        const double xp[3] = {(double)q[2], (double)q[3], 1.0};
        double Hx[3];
        for (int r = 0; r < 3; r++) Hx[r] = (Hd[r][0] * x0 + Hd[r][1] * y0) + Hd[r][2];
        cross3(Hx, xp, l[i]);
    }
    double e[3];
    cross3(l[0], l[1], e);
    double R[3][3];
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

int or_h_inlier_f32(const float H[9], const float* p, float t)
{
    const float t2 = t * t;
    const float z = (H[8] + H[6] * p[0]) + H[7] * p[1];
    const float z2 = z * z;
    const float d = fabsf(H[0] * H[4] - H[3] * H[1]);
    const float sc = d < z2 ? d : z2;
    const float u = ((H[2] + H[0] * p[0]) + H[1] * p[1]) - p[2] * z;
    const float v = ((H[5] + H[3] * p[0]) + H[4] * p[1]) - p[3] * z;
    return u * u + v * v < t2 * sc;
}

int or_fit_ls_f(const float* pts, const uint32_t* sel, uint32_t m, float F[9])
{
    return m >= 8 && fit_ls(pts, sel, m, 0, F);
}

int or_fit_ls_h(const float* pts, const uint32_t* sel, uint32_t m, float H[9])
{
    return m >= 4 && fit_ls(pts, sel, m, 1, H);
}

/* ---- finisher ---- */

typedef struct {
    const float* pts;
    uint32_t n;
    float t_f, t_h;
    int mode;
    uint8_t* bits;   /* inlier flags of the last counted model */
    uint8_t* planar; /* inlier flags of the current homography */
} dg_ctx;

static uint32_t count_f(dg_ctx* c, const float F[9], uint8_t* dst)
{
    uint32_t k = 0;
    for (uint32_t i = 0; i < c->n; i++) {
        dst[i] = (uint8_t)or_f_inlier_f32(F, c->pts + (size_t)i * 4, c->t_f, c->mode);
        k += dst[i];
    }
    return k;
}

static uint32_t count_h(dg_ctx* c, const float H[9], uint8_t* dst)
{
    uint32_t k = 0;
    for (uint32_t i = 0; i < c->n; i++) {
        dst[i] = (uint8_t)or_h_inlier_f32(H, c->pts + (size_t)i * 4, c->t_h);
        k += dst[i];
    }
    return k;
}

/* members of flags with rank r = floor(j total / 128), in j order */
static uint32_t select_evenly(const uint8_t* flags, uint32_t n, uint32_t total, uint32_t* sel)
{
    const int all = total <= DG_MAX_SEL;
    uint32_t r = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (!flags[i]) continue;
        if (all) sel[r] = i;
        else {
            const uint32_t j = (r * (uint32_t)DG_MAX_SEL + total - 1u) / total;
            if (j < DG_MAX_SEL && (j * total) / DG_MAX_SEL == r) sel[j] = i;
        }
        r++;
    }
    return all ? total : DG_MAX_SEL;
}

static uint32_t optimise_f(dg_ctx* c, float F[9])
{
    const uint32_t n0 = count_f(c, F, c->bits);
    if (n0 < 8) return n0;
    uint32_t sel[DG_MAX_SEL];
    const uint32_t m = select_evenly(c->bits, c->n, n0, sel);
    float F1[9];
    if (!fit_ls(c->pts, sel, m, 0, F1)) return n0;
    const uint32_t n1 = count_f(c, F1, c->bits);
    if (n1 <= n0) return n0;
    memcpy(F, F1, sizeof F1);
    return n1;
}

static uint32_t optimise_h(dg_ctx* c, float H[9])
{
    const uint32_t n0 = count_h(c, H, c->planar);
    if (n0 < 4) return n0;
    uint32_t sel[DG_MAX_SEL];
    const uint32_t m = select_evenly(c->planar, c->n, n0, sel);
    float H1[9];
    if (!fit_ls(c->pts, sel, m, 1, H1)) return n0;
    const uint32_t n1 = count_h(c, H1, c->bits);
    if (n1 <= n0) return n0;
    memcpy(H, H1, sizeof H1);
    memcpy(c->planar, c->bits, c->n);
    return n1;
}

static uint64_t splitmix(uint64_t* s)
{
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint32_t off_plane_at(const dg_ctx* c, uint32_t r)
{
    for (uint32_t i = 0; i < c->n; i++)
        if (!c->planar[i]) {
            if (r == 0) return i;
            r--;
        }
    return 0;
}

static int plane_sample(const dg_ctx* c, const float H[9], uint32_t m_off, uint64_t seed, uint32_t cand, uint32_t sidx,
                        float F[9])
{
    uint64_t st = seed ^ ((uint64_t)(cand * DG_MAX_PLANE + sidx) * 0xD1B54A32D192ED03ull);
    const uint32_t r0 = (uint32_t)(((splitmix(&st) >> 32) * (uint64_t)m_off) >> 32);
    uint32_t r1 = (uint32_t)(((splitmix(&st) >> 32) * (uint64_t)(m_off - 1u)) >> 32);
    if (r1 >= r0) r1++;
    const uint32_t ia = off_plane_at(c, r0), ib = off_plane_at(c, r1);
    return or_f_from_h(H, c->pts + (size_t)ia * 4, c->pts + (size_t)ib * 4, F);
}

void or_degensac_refine(const float* pts, uint32_t n, const uint32_t* idx, const uint32_t* counts, uint32_t nhyp,
                        float t_f, float t_h, int mode, uint32_t n_cand, uint32_t n_plane, uint64_t seed, float F_out[9],
                        float H_out[9], uint32_t result[4], uint8_t* mask_f, uint8_t* mask_h)
{
    dg_ctx c = {pts, n, t_f, t_h, mode, calloc(n ? n : 1, 1), calloc(n ? n : 1, 1)};
    float bestF[9] = {0}, bestH[9] = {0};
    uint32_t best_f = 0, best_h = 0, source = 0, best_hyp = ~0u;
    uint64_t prev = ~0ull;
    for (uint32_t cand = 0; n >= 7 && cand < n_cand; cand++) {
        uint64_t key = 0;
        for (uint32_t h = 0; h < nhyp; h++) {
            const uint64_t k = ((uint64_t)counts[h] << 32) | (uint32_t)~h;
            if (counts[h] > 0 && k < prev && k > key) key = k;
        }
        if (key == 0) break;
        prev = key;
        const uint32_t h = ~(uint32_t)key;
        int in_range = 1;
        for (int i = 0; i < 7; i++) in_range &= idx[(size_t)h * 7 + i] < n;
        float sample[7][4];
        for (int i = 0; i < 7; i++) memcpy(sample[i], pts + (size_t)(in_range ? idx[(size_t)h * 7 + i] : 0) * 4, 16);
        float models[27];
        const int nm = in_range ? or_f7_fit_f32(&sample[0][0], models) : 0;
        uint32_t cnt = 0;
        int bm = -1;
        for (int m = 0; m < nm; m++) {
            const uint32_t k = optimise_f(&c, models + 9 * m);
            if (k > cnt) {
                cnt = k;
                bm = m;
            }
        }
        if (bm < 0 || cnt <= best_f) continue;
        best_f = cnt;
        source = 0;
        best_hyp = h;
        memcpy(bestF, models + 9 * bm, sizeof bestF);
        uint32_t nb_planar = 3;
        float curH[9] = {0};
        for (int i = 0; i < 5; i++) {
            float tri[3][4], Hc[9];
            for (int k = 0; k < 3; k++) memcpy(tri[k], sample[kTri[i][k]], 16);
            if (!or_h_from_f(bestF, &tri[0][0], Hc)) continue;
            uint32_t k3 = 3;
            for (int k = 0; k < 4; k++) k3 += (uint32_t)or_h_inlier_f32(Hc, sample[kRem[i][k]], t_h);
            if (k3 > nb_planar) {
                nb_planar = k3;
                memcpy(curH, Hc, sizeof Hc);
            }
        }
        if (nb_planar < 5) continue;
        const uint32_t cnt_h = optimise_h(&c, curH);
        if (cnt_h > best_h) {
            best_h = cnt_h;
            memcpy(bestH, curH, sizeof bestH);
        }
        uint32_t m_off = 0;
        for (uint32_t i = 0; i < n; i++) m_off += !c.planar[i];
        if (m_off < 2) continue;
        uint64_t best = 0;
        for (uint32_t s = 0; s < n_plane; s++) {
            float Fh[9];
            if (!plane_sample(&c, curH, m_off, seed, cand, s, Fh)) continue;
            uint32_t k = 0;
            for (uint32_t i = 0; i < n; i++)
                k += (uint32_t)(!c.planar[i] && or_f_inlier_f32(Fh, pts + (size_t)i * 4, t_f, mode));
            const uint64_t key2 = ((uint64_t)k << 32) | (uint32_t)~s;
            if (key2 > best) best = key2;
        }
        if (best == 0) continue;
        float Fh[9];
        plane_sample(&c, curH, m_off, seed, cand, ~(uint32_t)best, Fh);
        const uint32_t k = optimise_f(&c, Fh);
        if (k > best_f) {
            best_f = k;
            source = 1;
            memcpy(bestF, Fh, sizeof bestF);
        }
    }
    memcpy(F_out, bestF, sizeof bestF);
    memcpy(H_out, bestH, sizeof bestH);
    result[0] = best_f;
    result[1] = best_h;
    result[2] = source;
    result[3] = best_hyp;
    for (uint32_t i = 0; i < n; i++) {
        if (mask_f) mask_f[i] = best_f > 0 ? (uint8_t)or_f_inlier_f32(bestF, pts + (size_t)i * 4, t_f, mode) : 0;
        if (mask_h) mask_h[i] = best_h > 0 ? (uint8_t)or_h_inlier_f32(bestH, pts + (size_t)i * 4, t_h) : 0;
    }
    free(c.bits);
    free(c.planar);
}

/* ---- the end of the 8-point F-RANSAC driver: adaptive cut and refit on all
 * inliers (ransac.hpp:139-160, ImagePair.cpp:185-203), restated for
 * rsfm_ransac_f8_finish ---- */

/* log2 of x in (0, 1] (0 -> -inf, negative -> NaN): x = m 2^e with m in [sqrt(1/2), sqrt(2)], ln m =
 * 2 atanh((m - 1) / (m + 1)) by its series to z^11 (z = s^2), evaluated as
 * s (2 + z P(z)) in Horner order. The GPU repeats every operation; the result
 * is pinned against glibc's log2 in tests. */
double or_det_log2(double x)
{
    if (x == 1.0) return 0.0;
    if (x == 0.0) return -INFINITY;
    if (!(x > 0.0)) return NAN; /* negative or NaN, as libm */
    uint64_t bits;
    memcpy(&bits, &x, 8);
    int e = (int)((bits >> 52) & 0x7FF) - 1023;
    bits = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    memcpy(&m, &bits, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    static const double coef[11] = {2.0 / 23.0, 2.0 / 21.0, 2.0 / 19.0, 2.0 / 17.0, 2.0 / 15.0, 2.0 / 13.0,
                                    2.0 / 11.0, 2.0 / 9.0,  2.0 / 7.0,  2.0 / 5.0,  2.0 / 3.0};
    double p = coef[0];
    for (int k = 1; k < 11; k++) p = p * z + coef[k];
    const double lnm = s * (2.0 + z * p);
    return (double)e + lnm * 1.4426950408889634;
}

/* ransac.hpp:155-157 with or_det_log2 and (best / n)^8 by three squarings.
 * Non-finite quotients convert as the reference built for x86-64 does:
 * -inf and NaN -> maxIter, +inf -> 0. */
uint32_t or_ransac_required_iters_det(uint32_t best, uint32_t n, float conf, uint32_t max_iter)
{
    const double num = or_det_log2(1.0 - (double)conf);
    const double p = (double)best / (double)n;
    const double p2 = p * p, p4 = p2 * p2, p8 = p4 * p4;
    const double q = num / or_det_log2(1.0 - p8);
    if (q != q || q == -INFINITY) return max_iter;
    if (q == INFINITY) return 0u;
    const double c = ceil(q);
    if (!(c > 0.0)) return 0u;
    if (c >= (double)max_iter) return max_iter;
    return (uint32_t)c;
}

/* the reference's loop (ransac.hpp:139-160) walking a scored table instead of
 * drawing: returns the hypotheses walked; best_h = ~0u when none was sane */
uint32_t or_ransac_adaptive_cut(const uint32_t* counts, uint32_t H, uint32_t n, float conf, uint32_t max_iter,
                                uint32_t* best_h, uint32_t* best_count)
{
    uint32_t best = 0, bh = ~0u, required = max_iter, iters = 0, insane = 0;
    const uint32_t cap = max_iter < 128u ? max_iter : 128u;
    uint32_t h = 0;
    while (h < H) {
        const uint32_t r = counts[h++];
        if (r == 0) {
            insane++;
            if (insane > cap) break;
            continue;
        }
        insane = 0;
        if (r > best) {
            best = r;
            bh = h - 1;
            required = or_ransac_required_iters_det(best, n, conf, max_iter);
        }
        iters++;
        if (iters >= required) break;
    }
    *best_h = bh;
    *best_count = best;
    return h;
}

#define FIN_THREADS 256u

/* an entry's 256 thread partials in the GPU's order: group g (0..15) adds
 * the partials of threads g, g + 16, ..., g + 240 in that order, then the 16
 * group sums are added in group order */
static double sum_partials(const double part[FIN_THREADS])
{
    double tot = 0.0;
    for (uint32_t g = 0; g < 16; g++) {
        double a = 0.0;
        for (uint32_t k = 0; k < 16; k++) a = a + part[k * 16 + g];
        tot = tot + a;
    }
    return tot;
}

/* least-squares F on the inliers of F0 (m of them) in the GPU's summation
 * order: thread t of 256 sums points t, t + 256, ... in index order, the 256
 * partials are added by sum_partials */
static int refit_all(const float* pts, uint32_t n, const float F0[9], float t, int mode, uint32_t m, float G[9])
{
    const double dm = (double)m;
    double part[45][FIN_THREADS];
    double tot[45];
    for (uint32_t th = 0; th < FIN_THREADS; th++) {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = th; i < n; i += FIN_THREADS) {
            const float* p = pts + (size_t)i * 4;
            if (!or_f_inlier_f32(F0, p, t, mode)) continue;
            for (int k = 0; k < 4; k++) a[k] = a[k] + (double)p[k];
        }
        for (int k = 0; k < 4; k++) part[k][th] = a[k];
    }
    for (int k = 0; k < 4; k++) tot[k] = sum_partials(part[k]);
    dg_norm nm;
    for (int k = 0; k < 4; k++) nm.c[k] = tot[k] / dm;
    for (uint32_t th = 0; th < FIN_THREADS; th++) {
        double d0 = 0.0, d1 = 0.0;
        for (uint32_t i = th; i < n; i += FIN_THREADS) {
            const float* p = pts + (size_t)i * 4;
            if (!or_f_inlier_f32(F0, p, t, mode)) continue;
            const double dx0 = (double)p[0] - nm.c[0], dy0 = (double)p[1] - nm.c[1];
            const double dx1 = (double)p[2] - nm.c[2], dy1 = (double)p[3] - nm.c[3];
            d0 = d0 + sqrt(dx0 * dx0 + dy0 * dy0);
            d1 = d1 + sqrt(dx1 * dx1 + dy1 * dy1);
        }
        part[0][th] = d0;
        part[1][th] = d1;
    }
    for (int k = 0; k < 2; k++) tot[k] = sum_partials(part[k]);
    const double n0 = tot[0] / dm, n1 = tot[1] / dm;
    if (!(n0 > 0.0) || !(n1 > 0.0)) return 0;
    nm.s0 = kSqrt2 / n0;
    nm.s1 = kSqrt2 / n1;
    for (uint32_t th = 0; th < FIN_THREADS; th++) {
        double acc[45];
        for (int e = 0; e < 45; e++) acc[e] = 0.0;
        for (uint32_t i = th; i < n; i += FIN_THREADS) {
            const float* p = pts + (size_t)i * 4;
            if (!or_f_inlier_f32(F0, p, t, mode)) continue;
            const double x0 = ((double)p[0] - nm.c[0]) * nm.s0, y0 = ((double)p[1] - nm.c[1]) * nm.s0;
            const double x1 = ((double)p[2] - nm.c[2]) * nm.s1, y1 = ((double)p[3] - nm.c[3]) * nm.s1;
            const double r[9] = {x1 * x0, y1 * x0, x0, x1 * y0, y1 * y0, y0, x1, y1, 1.0};
            int e = 0;
            for (int u = 0; u < 9; u++)
                for (int v = u; v < 9; v++) {
                    acc[e] = acc[e] + r[u] * r[v];
                    e++;
                }
        }
        for (int e = 0; e < 45; e++) part[e][th] = acc[e];
    }
    double N[9][9];
    int e = 0;
    for (int u = 0; u < 9; u++)
        for (int v = u; v < 9; v++) {
            const double a = sum_partials(part[e]);
            N[u][v] = a;
            N[v][u] = a;
            e++;
        }
    double v[9];
    return null_vector9(N, v) && finish_f(v, &nm, G);
}

static uint32_t count_inliers_f(const float* pts, uint32_t n, const float F[9], float t, int mode)
{
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; i++) k += (uint32_t)or_f_inlier_f32(F, pts + (size_t)i * 4, t, mode);
    return k;
}

int or_f8_finish(const float* pts, uint32_t n, const uint32_t* idx, const uint32_t* counts, const float* models,
                 uint32_t H, float t, int mode, float conf, uint32_t max_iter, int refit, float F[9],
                 uint32_t result[4], uint8_t* mask)
{
    if (n < 8) {
        memset(F, 0, 9 * sizeof(float));
        result[0] = ~0u;
        result[1] = result[2] = result[3] = 0u;
        if (mask) memset(mask, 0, n);
        return 0;
    }
    uint32_t bh = ~0u, bc = 0, walked = H;
    if (conf > 0.f) {
        walked = or_ransac_adaptive_cut(counts, H, n, conf, max_iter, &bh, &bc);
    } else {
        for (uint32_t h = 0; h < H; h++)
            if (counts[h] > bc) {
                bc = counts[h];
                bh = h;
            }
    }
    float F0[9];
    int ok;
    if (bh != ~0u && models) {
        memcpy(F0, models + (size_t)bh * 9, sizeof F0);
        ok = 1;
    } else {
        float p8[32];
        for (int i = 0; i < 8; i++) {
            const uint32_t k = bh != ~0u ? idx[(size_t)bh * 8 + i] : (uint32_t)i;
            memcpy(p8 + 4 * i, pts + (size_t)(k < n ? k : 0) * 4, 16);
        }
        ok = or_f8_fit_f32(p8, F0);
    }
    if (!ok) memset(F0, 0, sizeof F0);
    const uint32_t c0 = ok ? count_inliers_f(pts, n, F0, t, mode) : 0u;
    uint32_t cf = c0;
    if (refit && ok && c0 >= 8) {
        float G[9];
        if (refit_all(pts, n, F0, t, mode, c0, G)) {
            const uint32_t c1 = count_inliers_f(pts, n, G, t, mode);
            if (c1 > c0) {
                cf = c1;
                memcpy(F0, G, sizeof G);
            }
        }
    }
    memcpy(F, F0, sizeof F0);
    result[0] = bh;
    result[1] = c0;
    result[2] = walked;
    result[3] = cf;
    if (mask)
        for (uint32_t i = 0; i < n; i++) mask[i] = ok ? (uint8_t)or_f_inlier_f32(F0, pts + (size_t)i * 4, t, mode) : 0;
    return ok;
}
