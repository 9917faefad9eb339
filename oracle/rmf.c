/*
 * rmf.c - the per-cell affine RANSAC match filter restated on the CPU.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * Restates:
 *  - RansacMatchFilter::getInlierMask host binning  RansacMatchFilter.cpp:42-67, 130-133
 *    (cellWidth = ceil(max(w,h)/cellCols); cell = (ceil(y/cw), ceil(x/cw)) clamped)
 *  - kernelFindBestAffineForCells  RansacMatchFilter.cu:196-279
 *  - kernelTryBestAffineOfOtherCellsAndVote  RansacMatchFilter.cu:281-338
 *  - kernelComputeVotes  RansacMatchFilter.cu:341-391
 *  - CtaReducer::findMax tree (RansacMatchFilter.cu:160-180): thread t keeps its
 *    value unless thread t+mid holds a strictly larger count; simulated exactly.
 *  - solveGaussElim (no pivoting) with the operation order of SfmUtils.hpp:156-184.
 * Deliberate differences, each a determinism fix:
 *  - rngState = seed + gridTid*22695477 instead of clock() + ... (RansacMatchFilter.cu:216);
 *  - a cell with fewer than 9 neighbourhood matches writes a NaN affine (the
 *    reference leaves its scratch slot unwritten, so later kernels read stale memory);
 *  - voting uses the CTA-best count for every thread (the reference tests each
 *    thread's partial count before a __syncthreads, RansacMatchFilter.cu:326-328).
 * FP32 operation order (no fused multiply-add; build with -ffp-contract=off):
 *  diff_r = ((M_r0*lx + M_r1*ly) + t_r) - right_r ; sq = diff_0^2 + diff_1^2.
 * Pinning: the LCG against glibc rand_r; Gauss elimination against numpy's
 * solver; binning and votes on hand-built grids with known answers
 * (tests/test_oracle_rmf.py). The reference's kernel seeds its LCG from
 * clock() and uses cuda-utils kmat (absent), so its exact votes cannot be
 * reproduced.
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RMF_CTA 128
#define RMF_MIN_INLIERS 6

static inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

uint32_t or_rmf_bin(int img_w, int img_h, uint32_t cell_cols, const or_pt_pair* m, uint32_t n,
                    uint32_t* cols_out, uint32_t* rows_out, uint32_t* cw_out, or_pt_pair* reordered,
                    uint32_t* bounds, uint32_t* order)
{
    const uint32_t mx = (uint32_t)(img_w > img_h ? img_w : img_h);
    const uint32_t cw = div_up(mx, cell_cols);
    const uint32_t cols = div_up((uint32_t)img_w, cw);
    const uint32_t rows = div_up((uint32_t)img_h, cw);
    const uint32_t ncell = cols * rows;
    uint32_t* cell = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* cnt = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
    for (uint32_t k = 0; k < n; k++) {
        float fi = ceilf(m[k].ly / (float)cw);
        float fj = ceilf(m[k].lx / (float)cw);
        const float ri = (float)rows - 1.f, rj = (float)cols - 1.f;
        fi = fi < 0.f ? 0.f : (fi > ri ? ri : fi);
        fj = fj < 0.f ? 0.f : (fj > rj ? rj : fj);
        cell[k] = (uint32_t)fi * cols + (uint32_t)fj;
        cnt[cell[k] + 1]++;
    }
    for (uint32_t c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    memcpy(bounds, cnt, sizeof(uint32_t) * (ncell + 1));
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t p = cnt[cell[k]]++;
        reordered[p] = m[k];
        order[p] = k;
    }
    free(cnt);
    free(cell);
    *cols_out = cols;
    *rows_out = rows;
    *cw_out = cw;
    return ncell;
}

void or_solve_gauss3(const float A[9], const float b[6], float x[6])
{
    float Ab[3][5];
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) Ab[r][c] = A[r * 3 + c];
        Ab[r][3] = b[r * 2];
        Ab[r][4] = b[r * 2 + 1];
    }
    for (int i = 0; i < 3; i++) {
        const float inv = 1.f / Ab[i][i];
        for (int j = i + 1; j < 3; j++) {
            const float factor = Ab[j][i] * inv;
            for (int k = i + 1; k < 5; k++) Ab[j][k] = Ab[j][k] - factor * Ab[i][k];
        }
    }
    float X[3][2];
    for (int r = 0; r < 3; r++) { X[r][0] = Ab[r][3]; X[r][1] = Ab[r][4]; }
    for (int i = 2; i >= 0; i--) {
        const float inv = 1.f / Ab[i][i];
        X[i][0] = X[i][0] * inv;
        X[i][1] = X[i][1] * inv;
        for (int j = i - 1; j >= 0; j--) {
            const float factor = Ab[j][i];
            X[j][0] = X[j][0] - factor * X[i][0];
            X[j][1] = X[j][1] - factor * X[i][1];
        }
    }
    for (int r = 0; r < 3; r++) { x[r * 2] = X[r][0]; x[r * 2 + 1] = X[r][1]; }
}

typedef struct {
    uint32_t start[9], size[9], sum;
} cells_t;

static void cells_init(cells_t* c, const uint32_t* bounds, uint32_t cols, uint32_t rows, uint32_t bx, uint32_t by)
{
    c->sum = 0;
    for (int t = 0; t < 9; t++) {
        const int i = t / 3, j = t % 3;
        const int mm = (int)by - 1 + i, nn = (int)bx - 1 + j;
        if (mm >= 0 && mm < (int)rows && nn >= 0 && nn < (int)cols) {
            const uint32_t idx = (uint32_t)mm * cols + (uint32_t)nn;
            c->start[t] = bounds[idx];
            c->size[t] = bounds[idx + 1] - bounds[idx];
        } else {
            c->start[t] = 0;
            c->size[t] = 0;
        }
        c->sum += c->size[t];
    }
}

static uint32_t cells_map(const cells_t* c, uint32_t idx)
{
    uint32_t acc = 0;
    for (int i = 0; i < 9; i++) {
        if (c->size[i] == 0) continue;
        const uint32_t nxt = acc + c->size[i];
        if (idx < nxt) return c->start[i] + (idx - acc);
        acc = nxt;
    }
    return ~0u;
}

/* affine as 2x3 row-major: [m00 m01 t0; m10 m11 t1] */
static inline float sq_threshold(const float* T, float threshold)
{
    const float sqr_scale = T[0] * T[4] - T[1] * T[3];
    const float scale = sqrtf(sqr_scale);
    const float s = (1.f < scale) ? 1.f : scale; /* std::min(scale, 1.f) keeps NaN */
    const float th = threshold * s;
    return th * th;
}

static inline int is_inlier(const float* T, const or_pt_pair* p, float sq_thr)
{
    const float d0 = ((T[0] * p->lx + T[1] * p->ly) + T[2]) - p->rx;
    const float d1 = ((T[3] * p->lx + T[4] * p->ly) + T[5]) - p->ry;
    return d0 * d0 + d1 * d1 < sq_thr;
}

static uint32_t count_inliers(const cells_t* c, const float* T, const or_pt_pair* m, float threshold)
{
    const float sq = sq_threshold(T, threshold);
    uint32_t cnt = 0;
    for (int i = 0; i < 9; i++)
        for (uint32_t k = c->start[i]; k < c->start[i] + c->size[i]; k++) cnt += (uint32_t)is_inlier(T, &m[k], sq);
    return cnt;
}

/* CtaReducer::findMax over RMF_CTA per-thread (count, affine) */
static void cta_find_max(uint32_t* cnt, float (*aff)[6])
{
    for (uint32_t mid = RMF_CTA / 2; mid != 0; mid /= 2)
        for (uint32_t t = 0; t < mid; t++)
            if (cnt[t] < cnt[t + mid]) {
                cnt[t] = cnt[t + mid];
                memcpy(aff[t], aff[t + mid], sizeof(float) * 6);
            }
}

void or_rmf_filter(uint8_t* votes, float* affine, const or_pt_pair* m, uint32_t n,
                   const uint32_t* bounds, uint32_t cols, uint32_t rows, uint32_t nb_tests,
                   float threshold, int try_other, uint32_t seed)
{
    memset(votes, 0, n);
    if (n < 16) return;
    const uint32_t ncell = cols * rows;
    static const float ident[6] = {1, 0, 0, 0, 1, 0};
    /* kernel 1 */
    for (uint32_t by = 0; by < rows; by++)
        for (uint32_t bx = 0; bx < cols; bx++) {
            cells_t c;
            cells_init(&c, bounds, cols, rows, bx, by);
            const uint32_t rank = by * cols + bx;
            if (c.sum < 9) {
                for (int q = 0; q < 6; q++) affine[rank * 6 + q] = NAN;
                continue;
            }
            uint32_t tcnt[RMF_CTA];
            float taff[RMF_CTA][6];
            for (uint32_t tid = 0; tid < RMF_CTA; tid++) {
                tcnt[tid] = 0;
                memcpy(taff[tid], ident, sizeof(ident));
                const uint32_t grid_tid = RMF_CTA * rank + tid;
                int32_t st = (int32_t)(seed + grid_tid * 22695477u);
                for (uint32_t h = tid; h < nb_tests; h += RMF_CTA) {
                    uint32_t idx[3];
                    uint32_t got = 0;
                    while (got < 3) {
                        const uint32_t v = (uint32_t)or_lcg_step(&st) % c.sum;
                        int dup = 0;
                        for (uint32_t q = 0; q < got; q++) dup |= idx[q] == v;
                        if (!dup) idx[got++] = v;
                    }
                    float A[9], b[6], x[6];
                    for (int r = 0; r < 3; r++) {
                        const or_pt_pair* p = &m[cells_map(&c, idx[r])];
                        A[r * 3 + 0] = p->lx; A[r * 3 + 1] = p->ly; A[r * 3 + 2] = 1.f;
                        b[r * 2 + 0] = p->rx; b[r * 2 + 1] = p->ry;
                    }
                    or_solve_gauss3(A, b, x);
                    /* trans = x^T: row r = (x[0][r], x[1][r], x[2][r]) */
                    const float T[6] = {x[0], x[2], x[4], x[1], x[3], x[5]};
                    const float sqr_scale = T[0] * T[4] - T[1] * T[3];
                    if (sqr_scale < 0.0625f || sqr_scale > 16.f || !isfinite(sqr_scale)) continue;
                    const uint32_t k = count_inliers(&c, T, m, threshold);
                    if (k > tcnt[tid]) {
                        tcnt[tid] = k;
                        memcpy(taff[tid], T, sizeof(T));
                    }
                }
            }
            cta_find_max(tcnt, taff);
            memcpy(affine + rank * 6, taff[0], sizeof(float) * 6);
        }
    /* kernel 2 (vote) */
    for (uint32_t by = 0; by < rows; by++)
        for (uint32_t bx = 0; bx < cols; bx++) {
            cells_t c;
            cells_init(&c, bounds, cols, rows, bx, by);
            const uint32_t rank = by * cols + bx;
            float best[6];
            uint32_t bestc;
            if (try_other) {
                uint32_t tcnt[RMF_CTA];
                float taff[RMF_CTA][6];
                for (uint32_t tid = 0; tid < RMF_CTA; tid++) {
                    tcnt[tid] = 0;
                    memcpy(taff[tid], affine + rank * 6, sizeof(float) * 6);
                    for (uint32_t q = tid; q < ncell; q += RMF_CTA) {
                        const uint32_t k = count_inliers(&c, affine + q * 6, m, threshold);
                        if (k > tcnt[tid]) {
                            tcnt[tid] = k;
                            memcpy(taff[tid], affine + q * 6, sizeof(float) * 6);
                        }
                    }
                }
                cta_find_max(tcnt, taff);
                bestc = tcnt[0];
                memcpy(best, taff[0], sizeof(best));
            } else {
                memcpy(best, affine + rank * 6, sizeof(best));
                bestc = count_inliers(&c, best, m, threshold);
            }
            if (bestc < RMF_MIN_INLIERS) continue;
            const float sq = sq_threshold(best, threshold);
            for (uint32_t k = 0; k < c.sum; k++) {
                const uint32_t mi = cells_map(&c, k);
                if (is_inlier(best, &m[mi], sq)) votes[mi]++;
            }
        }
}
