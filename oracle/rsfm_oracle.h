/*
 * rsfm_oracle.h - CPU oracle for the two-view matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product path (rapidsfm_amd, librsfm_amd.so) never links or calls it.
 *
 * Each function restates the reference algorithm it names (file:line in the
 * lowsfer/RapidSFM tree). Pinning status of every function is listed in the
 * header comment of the .c file that implements it; functions whose result no
 * independent check fixes are marked "parity unpinned" there.
 */
#ifndef RSFM_ORACLE_H
#define RSFM_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- RNG (rng.c) ---------------- */
typedef struct { uint64_t mt[312]; int mti; } or_mt64;
void     or_mt64_seed(or_mt64* s, uint64_t seed);
uint64_t or_mt64_next(or_mt64* s);
/* uniform double in [0,1) from 53 high bits */
double   or_mt64_uniform(or_mt64* s);
/* standard normal via Box-Muller (cosine branch only, one draw per call) */
double   or_mt64_normal(or_mt64* s);
/* glibc TYPE_0 LCG step, as used by RansacMatchFilter.cu:45-50 */
int32_t  or_lcg_step(int32_t* state);

/* ---------------- synthetic inputs (gen.c) ---------------- */
/* SIFT-like u8 descriptors (values in [0,127]): half-normal, L2-normalise,
 * clip 0.2, renormalise, q = min(127, round(512 v)). Queries: n_true of them
 * are noisy copies of train descriptors (noise sigma added before the
 * normalise/clip/quantise chain), placed at shuffled positions; gt[q] is the
 * train index or -1. */
void or_gen_sift_pair(uint64_t seed, uint32_t n_train, uint32_t n_query, uint32_t n_true,
                      double noise_sigma, uint8_t* train, uint8_t* query, int32_t* gt);
/* uniformly random u8 descriptors over the full [0,255] range */
void or_gen_random_u8(uint64_t seed, uint32_t n, uint8_t* desc);

/* two-view geometry: n correspondences (x0,y0,x1,y1) float32, n_out of them
 * outliers uniform over the image; inlier flag per correspondence; the true
 * fundamental matrix F_true (row-major 3x3, double) with x1^T F x0 = 0. */
void or_gen_two_view(uint64_t seed, uint32_t n, uint32_t n_out, double focal, double width,
                     double height, double noise_px, float* pts, uint8_t* is_inlier, double* F_true);

/* one end-to-end image pair: descriptors (query A with n_true noisy copies of
 * train B) and keypoints (x,y) whose true pairs follow one two-view geometry
 * except an outlier_frac of them; gtA[i] = train index or -1. */
void or_gen_match_pair(uint64_t seed, uint32_t nA, uint32_t nB, uint32_t n_true, double outlier_frac,
                       double focal, double width, double height, double noise_px, uint8_t* descA,
                       uint8_t* descB, float* kpA, float* kpB, int32_t* gtA, uint8_t* geo_inlier);

/* sample index sets: H sets of k distinct indices in [0,n), non-duplicate
 * TableSampler semantics (ransac.hpp:52-82) driven by mt19937_64(seed). */
void or_sample_indices(uint64_t seed, uint32_t n, uint32_t k, uint32_t H, uint32_t* idx);
/* restatement of the GPU's splitmix64 sampler (rsfm_ransac_sample_indices) */
void or_sample_indices_hash(uint64_t seed, uint32_t n, uint32_t H, uint32_t* idx);
/* the same with k (4..8) indices per sample, rows of k */
void or_sample_indices_hash_k(uint64_t seed, uint32_t n, uint32_t k, uint32_t H, uint32_t* idx);
/* learned-proposal weighted draw (gen.c) */
void or_sample_weighted(const float* w, uint32_t n, uint32_t H, uint64_t seed, uint32_t* idx);

/* ---------------- brute-force matching (bf.c) ---------------- */
typedef struct { uint32_t index; uint32_t distance; } or_best_match;   /* BestMatchImpl<u32> */
typedef struct { uint32_t idxQuery; uint32_t idxTrain; float distance; } or_valid_match;

void or_sqrnorm_u8(const uint8_t* desc, uint32_t n, uint32_t* out);
void or_desc8b_to4b(const uint8_t* src, uint32_t n, uint8_t* dst);
void or_sqrnorm_u4(const uint8_t* desc4, uint32_t n, uint32_t* out);

/* exact top-1 bidirectional L2^2 match, lowest index on ties.
 * bestA/bestB initialised to {~0u, UINT32_MAX} first (empty side -> stays). */
void or_bf_match_u8(const uint8_t* A, uint32_t nA, const uint8_t* B, uint32_t nB,
                    or_best_match* bestA, or_best_match* bestB, int nthreads);
void or_bf_match_u4(const uint8_t* A4, uint32_t nA, const uint8_t* B4, uint32_t nB,
                    or_best_match* bestA, or_best_match* bestB, int nthreads);
/* top-2 for A (second best over j != best j), top-1 for B */
void or_bf_match_top2_u8(const uint8_t* A, uint32_t nA, const uint8_t* B, uint32_t nB,
                         or_best_match* bestA, or_best_match* secondA, or_best_match* bestB,
                         int nthreads);
/* same results, each distance computed once (CPU baseline path) */
void or_bf_match_top2_fast(const uint8_t* A, uint32_t nA, const uint8_t* B, uint32_t nB,
                           or_best_match* bestA, or_best_match* secondA, or_best_match* bestB, int nthreads);
/* dot-product ranking mode (higher is better, ties -> lowest index), top-2 */
void or_bf_match_dot_top2_u8(const uint8_t* A, uint32_t nA, const uint8_t* B, uint32_t nB,
                             int32_t* best_dot, uint32_t* best_idx, int32_t* second_dot,
                             uint32_t* second_idx, int nthreads);

uint32_t or_cross_check(const or_best_match* fwd, uint32_t nQ, const or_best_match* bwd,
                        uint32_t nT, or_valid_match* out);
uint32_t or_remove_conflicts(const or_best_match* fwd, uint32_t nQ, uint32_t nT, or_valid_match* out);
/* exact ratio test on L2^2 distances: sqrt(d1) < 0.8 sqrt(d2) <=> 25 d1 < 16 d2 */
int      or_ratio_pass(uint32_t d1, uint32_t d2, uint32_t num, uint32_t den);
/* cross-check + ratio test, ascending query index; returns count */
uint32_t or_filter_matches(const or_best_match* fwd, const or_best_match* second, uint32_t nQ,
                           const or_best_match* bwd, uint32_t nT, uint32_t ratio_num,
                           uint32_t ratio_den, or_valid_match* out);

/* ---------------- cascade hashing (cashash.c) ---------------- */
void or_cashash_accumulate(const uint8_t* desc, uint32_t n, uint64_t acc[128]);
/* nch = 64 in the reference; 192 extends it to the remap channels */
void or_cashash_compute_bias(const int8_t* w, const int32_t* bias_base, const uint64_t acc[128],
                             uint64_t nb_desc, uint32_t nch, int32_t* bias);
/* nch = 64 (bucket codes only) or 192 (+128-bit remap) */
void or_cashash_compute_hash(const int8_t* w, const int32_t* bias, uint32_t nch, const uint8_t* desc,
                             uint32_t n, uint8_t* codes /*n x 8*/, uint32_t* remap /*n x 4 or NULL*/);
void or_cashash_histogram(const uint8_t* codes, uint32_t n, uint32_t hist[8 * 256]);
uint32_t or_cashash_bucket_bounds(const uint32_t hist[8 * 256], uint32_t lower[8 * 256],
                                  uint32_t upper[8 * 256]);
/* buckets[s*n + slot] = desc index, ascending within each bucket (canonical) */
void or_cashash_build_buckets(const uint8_t* codes, uint32_t n, const uint32_t lower[8 * 256],
                              uint32_t* buckets);
/* union-of-buckets exact top-1 L2^2 both ways, nscheme tables used */
void or_cashash_find_best(const uint8_t* qdesc, const uint8_t* qcodes, uint32_t nq,
                          const uint8_t* tdesc, const uint8_t* tcodes, uint32_t nt,
                          uint32_t nscheme, or_best_match* bestQ, or_best_match* bestT);
/* six-table mode with Hamming re-rank: candidates = union over the first
 * nscheme tables; keep the `keep` candidates of least Hamming distance on the
 * 128-bit remap code (ties -> lower index), then exact L2^2 top-1 among them.
 * Query direction only. */
void or_cashash_find_best_rerank(const uint8_t* qdesc, const uint8_t* qcodes, const uint32_t* qremap,
                                 uint32_t nq, const uint8_t* tdesc, const uint8_t* tcodes,
                                 const uint32_t* tremap, uint32_t nt, uint32_t nscheme, uint32_t keep,
                                 or_best_match* bestQ);
/* weights/bias-base generator of bruteForceMatcher.cpp:127-141 is in
 * tests/golden (libstdc++ normal_distribution); this one is the seeded
 * Gaussian used for the six-table configuration. */
void or_cashash_gaussian_weights(uint64_t seed, uint32_t nch, int8_t* w, int32_t* bias_base);

/* ---------------- affine RANSAC match filter (rmf.c) ---------------- */
typedef struct { float lx, ly, rx, ry; } or_pt_pair;
/* host binning of RansacMatchFilter.cpp:42-67: returns cols*rows, fills
 * cols/rows/cellWidth, reordered matches, bounds (cells+1) and order. */
uint32_t or_rmf_bin(int img_w, int img_h, uint32_t cell_cols, const or_pt_pair* m, uint32_t n,
                    uint32_t* cols, uint32_t* rows, uint32_t* cell_width, or_pt_pair* reordered,
                    uint32_t* bounds, uint32_t* order);
/* kernels of RansacMatchFilter.cu with rngState = seed + gridTid*22695477 */
void or_rmf_filter(uint8_t* votes, float* affine /*cells x 6*/, const or_pt_pair* m, uint32_t n,
                   const uint32_t* bounds, uint32_t cols, uint32_t rows, uint32_t nb_tests,
                   float threshold, int try_other, uint32_t seed);
void or_solve_gauss3(const float A[9], const float b[6], float x[6]);

/* ---------------- F-matrix RANSAC (fransac.c) ---------------- */
/* FP32 8-point fit with the exact operation order of the GPU kernel.
 * pts8: 8 x (x0,y0,x1,y1). F out: row-major 3x3. Returns 0 when degenerate. */
int  or_f8_fit_f32(const float* pts8, float F[9]);
/* FP64 reference fit (Hartley + Jacobi SVD null vector + rank 2 by SVD) */
int  or_f8_fit_f64(const double* pts8, double F[9]);
/* inlier tests: mode 0 = reference two-sided line test (geometry.cpp:55-86),
 * mode 1 = Sampson, both threshold t in pixels */
int  or_f_inlier_f32(const float F[9], const float* p, float t, int mode);
int  or_f_inlier_f64(const double F[9], const double* p, double t, int mode);
/* score H hypotheses given sample indices; counts[h] (0 for rejected),
 * optional models (H x 9). min_sep_sq0/1: sample sanity gate (0 = off). */
void or_f8_score(const float* pts, uint32_t n, const uint32_t* idx, uint32_t H, float t, int mode,
                 float min_sep_sq0, float min_sep_sq1, uint32_t* counts, float* models, int nthreads);
/* full reference-style adaptive loop (ransac.hpp:115-163) with its sampler */
uint32_t or_f8_ransac_adaptive(const float* pts, uint32_t n, float t, int mode, float confidence,
                               uint32_t max_iter, uint64_t seed, float min_sep_sq0, float min_sep_sq1,
                               uint32_t best_idx[8], uint32_t* iters_out);
/* the same loop with score = ((fold of indices) * 2654435761) % modulus; used to
 * pin the loop against the reference's ransac<> compiled from ransac.hpp */
uint32_t or_ransac_loop_hash(uint32_t n, float confidence, uint32_t max_iter, uint64_t seed,
                             uint32_t modulus, uint32_t best_idx[8], uint32_t* draws_out);
uint32_t or_ransac_required_iters(uint32_t best, uint32_t n, uint32_t k, double conf, uint32_t max_iter);

/* ---------------- 7-point hypotheses (f7.c) ---------------- */
/* up to three F (row-major, unit Frobenius) from 7 correspondences; returns the count */
int  or_f7_fit_f32(const float* pts7 /*7 x 4*/, float* Fs /*3 x 9*/);
int  or_f7_fit_f64(const double* pts7, double* Fs);
/* per hypothesis ([H][7] indices): count and model of its best F */
void or_f7_score(const float* pts, uint32_t n, const uint32_t* idx, uint32_t H, float t, int mode, float min_sep_sq0,
                 float min_sep_sq1, uint32_t* counts, float* models, int nthreads);

/* ---------------- 4-point homography hypotheses (fransac.c) ---------------- */
int  or_h4_fit_f32(const float* pts4 /*4 x 4*/, float H[9]);
void or_h4_score(const float* pts, uint32_t n, const uint32_t* idx, uint32_t H, float t, float min_sep_sq0,
                 float min_sep_sq1, uint32_t* counts, float* models, int nthreads);

/* ---------------- relative pose from F (pose.c) ---------------- */
/* cameras as {fx, fy, cx, cy}; poses [4][12] = R row-major, t */
int  or_pose_decompose(const float F[9], const float c0[4], const float c1[4], float poses[48]);
int  or_pose_triangulate(const float P[12], const float* p, const float inv0[4], const float inv1[4], float X[3]);
void or_pose_from_f(const float F[9], const float c0[4], const float c1[4], const float* pts, uint32_t n,
                    const uint8_t* mask, float ratio, float poses[48], uint32_t result[10]);
/* homography branch: up to 8 valid candidates compacted, poses [8][12], normals [8][3] */
int  or_pose_decompose_h(const float H[9], const float c0[4], const float c1[4], float poses[96], float normals[24]);
void or_pose_from_h(const float H[9], const float c0[4], const float c1[4], const float* pts, uint32_t n,
                    const uint8_t* mask, float ratio, float poses[96], float normals[24], uint32_t result[19]);

/* ---------------- keypoint undistortion (undistort.c) ---------------- */
/* distortion model k = {k1, k2, p1, p2, k3} applied to normalised (x0, x1) */
void or_distort_f32(const float k[5], float x0, float x1, float out[2]);
void or_undistort_in_place(float* px, float* py, uint32_t n, const float cam[4], const float k[5], uint32_t iters);

/* ---------------- DEGENSAC finisher (degensac.c) ---------------- */
/* least-squares F (m >= 8) / H (m >= 4) on pts[sel[0..m)], unit Frobenius */
int  or_fit_ls_f(const float* pts, const uint32_t* sel, uint32_t m, float F[9]);
int  or_fit_ls_h(const float* pts, const uint32_t* sel, uint32_t m, float H[9]);
/* end of the 8-point F-RANSAC driver (ransac.hpp:139-160, ImagePair.cpp:185-203):
 * own FP64 log2, required iterations, the adaptive walk over a scored table and
 * rsfm_ransac_f8_finish restated (result = best, its count, walked, final count) */
double   or_det_log2(double x);
uint32_t or_ransac_required_iters_det(uint32_t best, uint32_t n, float conf, uint32_t max_iter);
uint32_t or_ransac_adaptive_cut(const uint32_t* counts, uint32_t H, uint32_t n, float conf, uint32_t max_iter,
                                uint32_t* best_h, uint32_t* best_count);
int      or_f8_finish(const float* pts, uint32_t n, const uint32_t* idx, const uint32_t* counts, const float* models,
                      uint32_t H, float t, int mode, float conf, uint32_t max_iter, int refit, float F[9],
                      uint32_t result[4], uint8_t* mask);
/* H compatible with F through three correspondences (3 x 4) */
int  or_h_from_f(const float F[9], const float* pts3, float H[9]);
/* F = [e']x H through two correspondences */
int  or_f_from_h(const float H[9], const float* pa, const float* pb, float F[9]);
/* checkProjectiveTransformation for one correspondence */
int  or_h_inlier_f32(const float H[9], const float* p, float t);
/* rsfm_degensac_refine for one pair */
void or_degensac_refine(const float* pts, uint32_t n, const uint32_t* idx, const uint32_t* counts, uint32_t nhyp,
                        float t_f, float t_h, int mode, uint32_t n_cand, uint32_t n_plane, uint64_t seed,
                        float F_out[9], float H_out[9], uint32_t result[4], uint8_t* mask_f, uint8_t* mask_h);

/* ---------------- learned proposal MLP (mlp.c) ---------------- */
/* 8 -> 64 -> 64 -> 64 -> 1, ReLU, sigmoid. params: W1[64x8] b1[64] W2[64x64]
 * b2[64] W3[64x64] b3[64] W4[64] b4[1] (float64 master copy). */
#define OR_MLP_NPARAM (64 * 8 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1)
void   or_mlp_forward(const double* params, const float* x /*n x 8*/, uint32_t n, double* prob);
/* mean binary cross-entropy loss and gradient over the batch */
double or_mlp_loss_grad(const double* params, const float* x, const float* y, uint32_t n,
                        double* grad);
/* the same with the GPU kernels' bf16 roundings emulated (mlp.c header) */
void   or_mlp_forward_bf16(const double* params, const float* x, uint32_t n, double* prob);
double or_mlp_loss_grad_bf16(const double* params, const float* x, const float* y, uint32_t n,
                             double* grad);
/* per-match features of the proposal network (rsfm_proposal_features_batch) */
void   or_proposal_features(const or_valid_match* m, uint32_t n, const or_best_match* second, const float* kpQ,
                            const float* kpT, float width, float height, float* x);
/* float -> bf16 -> float, round to nearest even */
double or_bf16_round(double v);

#ifdef __cplusplus
}
#endif
#endif
