/*
 * gen.c - seeded synthetic inputs shared by the tests, the benchmark and the
 * CPU baseline. TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * All draws come from MT19937-64 (rng.c) so the data is identical on any
 * toolchain. The reference has no generator for these workloads: its tests
 * draw random 32-bit words (test/bruteForceMatchKernelTest.cpp:31-35); the
 * SIFT-like rule and the two-view geometry below are this project's own.
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* normalise, clip at 0.2, renormalise, quantise: q = min(127, round(512 v)) */
static void sift_finish(double* v, uint8_t* out)
{
    double s = 0;
    for (int k = 0; k < 128; k++) s += v[k] * v[k];
    s = sqrt(s);
    if (s <= 0) s = 1;
    for (int k = 0; k < 128; k++) {
        v[k] /= s;
        if (v[k] > 0.2) v[k] = 0.2;
    }
    s = 0;
    for (int k = 0; k < 128; k++) s += v[k] * v[k];
    s = sqrt(s);
    if (s <= 0) s = 1;
    for (int k = 0; k < 128; k++) {
        v[k] /= s;
        double q = floor(512.0 * v[k] + 0.5);
        if (q > 127) q = 127;
        if (q < 0) q = 0;
        out[k] = (uint8_t)q;
    }
}

void or_gen_sift_pair(uint64_t seed, uint32_t n_train, uint32_t n_query, uint32_t n_true,
                      double noise_sigma, uint8_t* train, uint8_t* query, int32_t* gt)
{
    or_mt64 rng;
    or_mt64_seed(&rng, seed);
    double* tv = (double*)malloc(sizeof(double) * 128 * (size_t)(n_train ? n_train : 1));
    double v[128];
    for (uint32_t i = 0; i < n_train; i++) {
        for (int k = 0; k < 128; k++) v[k] = fabs(or_mt64_normal(&rng));
        sift_finish(v, train + (size_t)i * 128);
        /* keep the renormalised float vector for the noisy copies */
        memcpy(tv + (size_t)i * 128, v, sizeof(v));
    }
    if (n_true > n_query) n_true = n_query;
    if (n_true > n_train) n_true = n_train;
    /* which train descriptors get copied: first n_true of a Fisher-Yates permutation */
    uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (n_train ? n_train : 1));
    for (uint32_t i = 0; i < n_train; i++) perm[i] = i;
    for (uint32_t i = n_train; i > 1; i--) {
        const uint32_t j = (uint32_t)(or_mt64_next(&rng) % i);
        const uint32_t t = perm[i - 1]; perm[i - 1] = perm[j]; perm[j] = t;
    }
    /* query slot order */
    uint32_t* slot = (uint32_t*)malloc(sizeof(uint32_t) * (n_query ? n_query : 1));
    for (uint32_t i = 0; i < n_query; i++) slot[i] = i;
    for (uint32_t i = n_query; i > 1; i--) {
        const uint32_t j = (uint32_t)(or_mt64_next(&rng) % i);
        const uint32_t t = slot[i - 1]; slot[i - 1] = slot[j]; slot[j] = t;
    }
    for (uint32_t q = 0; q < n_query; q++) {
        const uint32_t dst = slot[q];
        if (q < n_true) {
            const double* src = tv + (size_t)perm[q] * 128;
            for (int k = 0; k < 128; k++) {
                const double x = src[k] + noise_sigma * or_mt64_normal(&rng);
                v[k] = x > 0 ? x : 0;
            }
            if (gt) gt[dst] = (int32_t)perm[q];
        } else {
            for (int k = 0; k < 128; k++) v[k] = fabs(or_mt64_normal(&rng));
            if (gt) gt[dst] = -1;
        }
        sift_finish(v, query + (size_t)dst * 128);
    }
    free(slot);
    free(perm);
    free(tv);
}

void or_gen_random_u8(uint64_t seed, uint32_t n, uint8_t* desc)
{
    or_mt64 rng;
    or_mt64_seed(&rng, seed);
    for (size_t i = 0; i < (size_t)n * 128; i += 8) {
        const uint64_t r = or_mt64_next(&rng);
        memcpy(desc + i, &r, 8);
    }
}

/* Rodrigues rotation from axis-angle */
static void rodrigues(const double w[3], double R[9])
{
    const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (th < 1e-15) {
        for (int i = 0; i < 9; i++) R[i] = (i % 4 == 0) ? 1 : 0;
        return;
    }
    const double k[3] = {w[0] / th, w[1] / th, w[2] / th};
    const double c = cos(th), s = sin(th), C = 1 - c;
    R[0] = c + k[0] * k[0] * C;        R[1] = k[0] * k[1] * C - k[2] * s; R[2] = k[0] * k[2] * C + k[1] * s;
    R[3] = k[1] * k[0] * C + k[2] * s; R[4] = c + k[1] * k[1] * C;        R[5] = k[1] * k[2] * C - k[0] * s;
    R[6] = k[2] * k[0] * C - k[1] * s; R[7] = k[2] * k[1] * C + k[0] * s; R[8] = c + k[2] * k[2] * C;
}

static void mat3_mul(const double* A, const double* B, double* C)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += A[i * 3 + k] * B[k * 3 + j];
            C[i * 3 + j] = s;
        }
}

void or_gen_two_view(uint64_t seed, uint32_t n, uint32_t n_out, double focal, double width,
                     double height, double noise_px, float* pts, uint8_t* is_inlier, double* F_true)
{
    or_mt64 rng;
    or_mt64_seed(&rng, seed);
    const double cx = width / 2, cy = height / 2;
    /* camera 2: rotation angle <= 15 deg about a random axis, baseline 1..3 */
    double axis[3], t[3];
    double an = 0;
    do {
        for (int i = 0; i < 3; i++) axis[i] = or_mt64_normal(&rng);
        an = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    } while (an < 1e-6);
    const double ang = (15.0 * M_PI / 180.0) * or_mt64_uniform(&rng);
    double w[3] = {axis[0] / an * ang, axis[1] / an * ang, axis[2] / an * ang};
    double R[9];
    rodrigues(w, R);
    double tn = 0;
    do {
        for (int i = 0; i < 3; i++) t[i] = or_mt64_normal(&rng);
        tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    } while (tn < 1e-6);
    const double base = 1.0 + 2.0 * or_mt64_uniform(&rng);
    for (int i = 0; i < 3; i++) t[i] = t[i] / tn * base;
    /* F = K^-T [t]x R K^-1 ; x1^T F x0 = 0 with X1 = R X0 + t */
    if (F_true) {
        const double tx[9] = {0, -t[2], t[1], t[2], 0, -t[0], -t[1], t[0], 0};
        double E[9];
        mat3_mul(tx, R, E);
        const double Ki[9] = {1 / focal, 0, -cx / focal, 0, 1 / focal, -cy / focal, 0, 0, 1};
        double KiT[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) KiT[i * 3 + j] = Ki[j * 3 + i];
        double T1[9];
        mat3_mul(KiT, E, T1);
        mat3_mul(T1, Ki, F_true);
    }
    /* which correspondences are outliers: first n_out of a permutation */
    uint8_t* outl = (uint8_t*)calloc(n ? n : 1, 1);
    uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    for (uint32_t i = n; i > 1; i--) {
        const uint32_t j = (uint32_t)(or_mt64_next(&rng) % i);
        const uint32_t tt = perm[i - 1]; perm[i - 1] = perm[j]; perm[j] = tt;
    }
    for (uint32_t i = 0; i < n_out && i < n; i++) outl[perm[i]] = 1;
    for (uint32_t i = 0; i < n; i++) {
        float* p = pts + (size_t)i * 4;
        if (outl[i]) {
            p[0] = (float)(width * or_mt64_uniform(&rng));
            p[1] = (float)(height * or_mt64_uniform(&rng));
            p[2] = (float)(width * or_mt64_uniform(&rng));
            p[3] = (float)(height * or_mt64_uniform(&rng));
            if (is_inlier) is_inlier[i] = 0;
            continue;
        }
        for (;;) {
            const double u = width * or_mt64_uniform(&rng);
            const double v = height * or_mt64_uniform(&rng);
            const double depth = 10.0 + 40.0 * or_mt64_uniform(&rng);
            const double X[3] = {(u - cx) / focal * depth, (v - cy) / focal * depth, depth};
            double Y[3];
            for (int r = 0; r < 3; r++) Y[r] = R[r * 3] * X[0] + R[r * 3 + 1] * X[1] + R[r * 3 + 2] * X[2] + t[r];
            if (Y[2] <= 1e-3) continue;
            const double u1 = focal * Y[0] / Y[2] + cx;
            const double v1 = focal * Y[1] / Y[2] + cy;
            if (u1 < 0 || u1 >= width || v1 < 0 || v1 >= height) continue;
            p[0] = (float)(u + noise_px * or_mt64_normal(&rng));
            p[1] = (float)(v + noise_px * or_mt64_normal(&rng));
            p[2] = (float)(u1 + noise_px * or_mt64_normal(&rng));
            p[3] = (float)(v1 + noise_px * or_mt64_normal(&rng));
            break;
        }
        if (is_inlier) is_inlier[i] = 1;
    }
    free(perm);
    free(outl);
}

/* one image pair for the end-to-end workload: descriptors as or_gen_sift_pair
 * (query = A, train = B, n_true noisy copies), keypoints from the two-view
 * geometry of or_gen_two_view: a fraction (1 - outlier_frac) of the true
 * descriptor pairs are projections of one 3-D point (+ noise_px), the other
 * keypoints are uniform over the image. kp arrays are (x, y) float32. */
void or_gen_match_pair(uint64_t seed, uint32_t nA, uint32_t nB, uint32_t n_true, double outlier_frac,
                       double focal, double width, double height, double noise_px, uint8_t* descA,
                       uint8_t* descB, float* kpA, float* kpB, int32_t* gtA, uint8_t* geo_inlier)
{
    or_gen_sift_pair(seed, nB, nA, n_true, 0.05, descB, descA, gtA);
    or_mt64 rng;
    or_mt64_seed(&rng, seed ^ 0x5DEECE66DULL);
    const double cx = width / 2, cy = height / 2;
    double axis[3], t[3], an, tn;
    do {
        for (int i = 0; i < 3; i++) axis[i] = or_mt64_normal(&rng);
        an = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    } while (an < 1e-6);
    const double ang = (15.0 * M_PI / 180.0) * or_mt64_uniform(&rng);
    const double w[3] = {axis[0] / an * ang, axis[1] / an * ang, axis[2] / an * ang};
    double R[9];
    rodrigues(w, R);
    do {
        for (int i = 0; i < 3; i++) t[i] = or_mt64_normal(&rng);
        tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    } while (tn < 1e-6);
    const double base = 1.0 + 2.0 * or_mt64_uniform(&rng);
    for (int i = 0; i < 3; i++) t[i] = t[i] / tn * base;
    for (uint32_t j = 0; j < nB; j++) {
        kpB[2 * j] = (float)(width * or_mt64_uniform(&rng));
        kpB[2 * j + 1] = (float)(height * or_mt64_uniform(&rng));
    }
    for (uint32_t i = 0; i < nA; i++) {
        if (geo_inlier) geo_inlier[i] = 0;
        const int32_t g = gtA[i];
        const int geo = g >= 0 && or_mt64_uniform(&rng) >= outlier_frac;
        if (!geo) {
            kpA[2 * i] = (float)(width * or_mt64_uniform(&rng));
            kpA[2 * i + 1] = (float)(height * or_mt64_uniform(&rng));
            continue;
        }
        for (;;) {
            const double u = width * or_mt64_uniform(&rng);
            const double v = height * or_mt64_uniform(&rng);
            const double depth = 10.0 + 40.0 * or_mt64_uniform(&rng);
            const double X[3] = {(u - cx) / focal * depth, (v - cy) / focal * depth, depth};
            double Y[3];
            for (int r = 0; r < 3; r++) Y[r] = R[r * 3] * X[0] + R[r * 3 + 1] * X[1] + R[r * 3 + 2] * X[2] + t[r];
            if (Y[2] <= 1e-3) continue;
            const double u1 = focal * Y[0] / Y[2] + cx;
            const double v1 = focal * Y[1] / Y[2] + cy;
            if (u1 < 0 || u1 >= width || v1 < 0 || v1 >= height) continue;
            kpA[2 * i] = (float)(u + noise_px * or_mt64_normal(&rng));
            kpA[2 * i + 1] = (float)(v + noise_px * or_mt64_normal(&rng));
            kpB[2 * g] = (float)(u1 + noise_px * or_mt64_normal(&rng));
            kpB[2 * g + 1] = (float)(v1 + noise_px * or_mt64_normal(&rng));
            break;
        }
        if (geo_inlier) geo_inlier[i] = 1;
    }
}

/* TableSampler::operator() of ransac.hpp:59-72 with Picker (ransac.hpp:31-38,
 * no focused tail): val = engine() % n, redraw while already taken. */
void or_sample_indices(uint64_t seed, uint32_t n, uint32_t k, uint32_t H, uint32_t* idx)
{
    or_mt64 rng;
    or_mt64_seed(&rng, seed);
    uint8_t* mark = (uint8_t*)calloc(n ? n : 1, 1);
    for (uint32_t h = 0; h < H; h++) {
        uint32_t* out = idx + (size_t)h * k;
        for (uint32_t i = 0; i < k; i++) {
            uint32_t val;
            do {
                val = (uint32_t)(or_mt64_next(&rng) % n);
            } while (mark[val]);
            out[i] = val;
            mark[val] = 1;
        }
        for (uint32_t i = 0; i < k; i++) mark[out[i]] = 0;
    }
    free(mark);
}

/* the device sampler of rsfm_ransac_sample_indices: hypothesis h draws from
 * splitmix64 seeded with seed ^ h * 0xD1B54A32D192ED03, value % n, repeats
 * rejected; n < 8 gives ~0u entries */
static uint64_t splitmix_next(uint64_t* s)
{
    *s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void or_sample_indices_hash_k(uint64_t seed, uint32_t n, uint32_t k, uint32_t H, uint32_t* idx)
{
    for (uint32_t h = 0; h < H; h++) {
        uint32_t* out = idx + (size_t)h * k;
        if (n < k) {
            for (uint32_t i = 0; i < k; i++) out[i] = ~0u;
            continue;
        }
        uint64_t s = seed ^ ((uint64_t)h * 0xD1B54A32D192ED03ULL);
        for (uint32_t i = 0; i < k; i++) {
            for (;;) {
                const uint32_t c = (uint32_t)(splitmix_next(&s) % n);
                int dup = 0;
                for (uint32_t j = 0; j < i; j++) dup |= out[j] == c;
                if (!dup) { out[i] = c; break; }
            }
        }
    }
}

void or_sample_indices_hash(uint64_t seed, uint32_t n, uint32_t H, uint32_t* idx)
{
    or_sample_indices_hash_k(seed, n, 8, H, idx);
}

/* Weighted draw of 8 distinct indices per hypothesis (the learned-proposal
 * sampler; no reference counterpart). Integer weights q_i = round(clamp(w_i,
 * 0, 1) * 65536), exact u64 prefix; r = splitmix % total picks the first i
 * with cdf[i] > r; repeats rejected, 512 draws at most (then all ~0u);
 * total 0 falls back to the uniform draw of or_sample_indices_hash. */
void or_sample_weighted(const float* w, uint32_t n, uint32_t H, uint64_t seed, uint32_t* idx)
{
    uint64_t* cdf = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        float v = w[i];
        v = v > 0.f ? v : 0.f; /* NaN -> 0 */
        v = v < 1.f ? v : 1.f;
        acc += (uint32_t)(v * 65536.f + 0.5f);
        cdf[i] = acc;
    }
    for (uint32_t h = 0; h < H; h++) {
        uint32_t* out = idx + (size_t)h * 8;
        if (n < 8) {
            for (int i = 0; i < 8; i++) out[i] = ~0u;
            continue;
        }
        const uint64_t total = cdf[n - 1];
        uint64_t s = seed ^ ((uint64_t)h * 0xD1B54A32D192ED03ULL);
        uint32_t draws = 0;
        int failed = 0;
        for (int i = 0; i < 8 && !failed; i++) {
            for (;;) {
                uint32_t c;
                if (total == 0) {
                    c = (uint32_t)(splitmix_next(&s) % n);
                } else {
                    const uint64_t r = splitmix_next(&s) % total;
                    uint32_t lo = 0, hi = n - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (cdf[mid] > r) hi = mid;
                        else lo = mid + 1;
                    }
                    c = lo;
                }
                int dup = 0;
                for (int j = 0; j < i; j++) dup |= out[j] == c;
                if (!dup) { out[i] = c; break; }
                if (++draws >= 512) { failed = 1; break; }
            }
        }
        if (failed)
            for (int i = 0; i < 8; i++) out[i] = ~0u;
    }
    free(cdf);
}
