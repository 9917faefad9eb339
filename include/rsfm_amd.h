/*
 * rsfm_amd.h - C ABI of the MI355X (gfx950) two-view matching hot path.
 *
 * Every entry point enqueues work on the caller's HIP stream and returns an
 * int status (RSFM_OK = 0, negative on error; rsfm_last_error() gives the
 * message of the calling thread's last failure). Where the reference throws
 * (cudaCheck -> cudapp::Exception) or asserts, these return an error code; the
 * C++ adapters in rsfm_amd.hpp turn a nonzero status back into an exception.
 * The caller owns every buffer. Descriptors are 128 B (u8 x 128, AoS) or 64 B
 * in the 4-bit mode; BestMatch is 8 B {u32 index; u32 distance}, index in the
 * low word, exactly BestMatchImpl<uint32_t> of the reference.
 * No entry point allocates device memory; the ones that need scratch take a
 * caller workspace and have a *_workspace_size query.
 */
#ifndef RSFM_AMD_H
#define RSFM_AMD_H
#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSFM_OK 0
#define RSFM_ERR_INVALID (-1)
#define RSFM_ERR_HIP (-2)
#define RSFM_ERR_UNSUPPORTED (-3)

/* message of this thread's last failed call ("" if none) */
const char* rsfm_last_error(void);
/* library version string */
const char* rsfm_version(void);

/* ---------------------------------------------------------------------------
 * Types (identical layouts to the reference)
 * ------------------------------------------------------------------------ */

/* BestMatchImpl<uint32_t>, bruteForceMatch.h:91-96 */
typedef struct rsfm_best_match {
    uint32_t index;
    uint32_t distance;
} rsfm_best_match;

/* BruteForceMatchTask<kU8|kU4,128,kL2>, bruteForceMatch.h:98-113 (56 B).
 * sqrNormDescA/B are accepted for layout compatibility: the kernels derive
 * the distance offsets from the descriptor bytes themselves, so results equal
 * the reference's whenever the norms passed are the true sums of squares. */
typedef struct rsfm_bf_task {
    uint32_t nbDescA;
    uint32_t nbDescB;
    const void* descA;
    const void* descB;
    rsfm_best_match* bestMatchA;
    rsfm_best_match* bestMatchB; /* may be NULL: query direction only */
    const uint32_t* sqrNormDescA;
    const uint32_t* sqrNormDescB;
} rsfm_bf_task;

/* ValidMatch, bruteForceMatch.h:120-125 */
typedef struct rsfm_valid_match {
    uint32_t idxQuery;
    uint32_t idxTrain;
    float distance;
} rsfm_valid_match;

/* top-2 task: no reference counterpart. bestMatchA/secondMatchA get the best
 * and second-best train index per query (second = best over j != best j,
 * lexicographic (L2^2, index)); bestMatchB the best query per train. */
typedef struct rsfm_bf_top2_task {
    rsfm_bf_task base;
    rsfm_best_match* secondMatchA;
} rsfm_bf_top2_task;

/* ---------------------------------------------------------------------------
 * Brute-force matching
 * ------------------------------------------------------------------------ */

/* replaces rsfm::launchPreCompSiftSqrNorm (bruteForceMatch.h:135,
 * bruteForceMatchIMMA.cu:743-750): sqrNorm[i] = sum_k desc[i][k]^2 */
int rsfm_sqrnorm_u8(uint32_t* sqrNorm, const void* desc, uint32_t nbDesc, hipStream_t stream);

/* replaces rsfm::launchBruteForceMatchIMMA(const SiftBruteForceMatchTask*, ...)
 * (bruteForceMatch.h:136, bruteForceMatchIMMA.cu:626-633): bidirectional top-1
 * L2^2, lowest index on ties; outputs start at {~0u, UINT32_MAX}. Any number
 * of tasks (internally batched). The train sets' staging form (132 B per
 * descriptor) comes from hipMallocAsync / hipFreeAsync on the stream. */
int rsfm_bf_match_u8(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream);

/* replaces rsfm::launchBruteForceMatchIDP (bruteForceMatch.h:118,
 * bruteForceMatch.cu:615-620): same contract on the vector ALUs (v_dot4),
 * no matrix cores. */
int rsfm_bf_match_idp_u8(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream);

/* replaces rsfm::launchSiftDesc8bTo4b (bruteForceMatch.h:133,
 * bruteForceMatchIMMA.cu:663-667): keep the high nibble of each byte,
 * dst word i = ((src[2i] & 0xF0F0F0F0) >> 4) | (src[2i+1] & 0xF0F0F0F0) */
int rsfm_desc8b_to_4b(void* dst, const void* src, uint32_t nbDesc, hipStream_t stream);

/* replaces rsfm::launchPreCompSift4bSqrNorm (bruteForceMatch.h:134) */
int rsfm_sqrnorm_u4(uint32_t* sqrNorm, const void* desc4, uint32_t nbDesc, hipStream_t stream);

/* replaces rsfm::launchBruteForceMatchIMMA(const Sift4bBruteForceMatchTask*, ...)
 * (bruteForceMatch.h:134): descA/descB point to 64-B 4-bit descriptors; the
 * staging form as for rsfm_bf_match_u8 */
int rsfm_bf_match_u4(const rsfm_bf_task* tasks, size_t nbTasks, hipStream_t stream);

/* top-2 query side + top-1 train side (new). Needs a workspace of
 * rsfm_bf_match_top2_workspace_size() bytes, 256-B aligned: the staging form
 * of each task's train descriptors (132 B per descriptor, nbDescB rounded up
 * to 64), written by the call, and the segment partials of long train sets. */
int rsfm_bf_match_top2_workspace_size(const rsfm_bf_top2_task* tasks, size_t nbTasks, size_t* bytes);
int rsfm_bf_match_top2(const rsfm_bf_top2_task* tasks, size_t nbTasks, void* workspace,
                       size_t workspaceBytes, hipStream_t stream);
/* rsfm_bf_match_top2 without a caller workspace and without any allocation (no
 * reference counterpart; for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops): the workspace lives in a fixed 64 MiB device
 * array of the library, held on every GPU the library runs on until the process
 * ends. Tasks needing more than 64 MiB of workspace (rsfm_bf_match_top2_workspace_size)
 * are rejected with RSFM_ERR_INVALID. Calls that overlap on one GPU race on the
 * array: serialise them (one stream, or events). */
int rsfm_bf_match_top2_noalloc(const rsfm_bf_top2_task* tasks, size_t nbTasks, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Cascade hashing (rsfm::cashash, casHash/cascadeHash.h:32-107)
 * Weights are int8 [nbChannels][128] (the reference's KArray<int32_t,64,32>
 * bytes); codes ("bucketIndices") are u8[8] per descriptor, code byte s =
 * channels 8s..8s+7 (bit b = channel 8s+b); remap is u32[4] per descriptor,
 * word r bit b = channel 64+32r+b. Histograms and bounds are u32[8][256].
 * ------------------------------------------------------------------------ */

/* replaces cashash::accumuateDesc (accumulateDesc.cu:117-131): acc[k] (+)=
 * sum over descriptors of (byte_k >> 1); sqrAcc (optional) the squares */
int rsfm_cashash_accumulate_desc(uint64_t* acc, uint64_t* sqrAcc, int initAcc, const void* desc, uint32_t nbDesc,
                                 hipStream_t stream);
/* replaces cashash::computeBias (accumulateDesc.cu:176-182): bias[c] =
 * int(roundf(sum_k w[c][k] * (-(float)acc[k] / (float)nbDesc))) + biasBase[c],
 * FP32 with fused multiply-add in k order per 16-term lane and an xor tree over
 * 8 lanes. nbChannels = 64 (reference) or 192 (bucket + remap channels). */
int rsfm_cashash_compute_bias(int32_t* bias, const int8_t* weights, const int32_t* biasBase, const uint64_t* acc,
                              uint64_t nbDesc, uint32_t nbChannels, hipStream_t stream);
/* replaces cashash::computeHashForBucketOnly (computeHash.cu:498-520): 64 channels;
 * histogram (optional) is cleared and filled */
int rsfm_cashash_compute_hash_bucket_only(const int8_t* weights, const int32_t* bias, const void* desc, uint32_t nbDesc,
                                          uint8_t* bucketIndices, uint32_t* histogram, hipStream_t stream);
/* replaces cashash::computeHashForBucketAndRemap (computeHash.cu:474-496): 192 channels */
int rsfm_cashash_compute_hash_bucket_and_remap(const int8_t* weights, const int32_t* bias, const void* desc,
                                               uint32_t nbDesc, uint8_t* bucketIndices, uint32_t* remap,
                                               uint32_t* histogram, hipStream_t stream);
/* replaces cashash::buildHistogram (buildBuckets.cu:107-114) */
int rsfm_cashash_build_histogram(uint32_t* histogram, const uint8_t* bucketIndices, uint32_t nbDesc, hipStream_t stream);
/* replaces cashash::computeBucketBound (buildBuckets.cu:169-172): lower =
 * exclusive, upper = inclusive prefix per scheme; *maxBucketSize (device,
 * optional) the largest bucket */
int rsfm_cashash_compute_bucket_bound(uint32_t* lower, uint32_t* upper, uint32_t* maxBucketSize,
                                      const uint32_t* histogram, hipStream_t stream);
/* replaces cashash::buildBuckets (buildBuckets.cu:262-269): buckets[s][slot]
 * = descriptor index; idxAllocator starts as `lower` and is consumed; order
 * inside a bucket is unspecified, as in the reference */
int rsfm_cashash_build_buckets(const uint8_t* bucketIndices, uint32_t nbDesc, uint32_t* idxAllocator, uint32_t* buckets,
                               hipStream_t stream);

/* cashash::ImageDesc (cascadeHash.h:60-92), 32 B */
typedef struct rsfm_cashash_image {
    const void* desc;
    const uint32_t* bucketEnds; /* [8][256] inclusive prefix (upper bounds) */
    const uint32_t* buckets;    /* [8][nbDesc] */
    const uint32_t* sqrNormDesc;
} rsfm_cashash_image;
/* cashash::MatchResult (cascadeHash.h:95-99), 16 B */
typedef struct rsfm_cashash_result {
    rsfm_best_match* query; /* indices into train */
    rsfm_best_match* train; /* indices into query, may be NULL */
} rsfm_cashash_result;
/* replaces cashash::findBestMatch (findBestMatch.cu:654-676): for every query
 * descriptor the lexicographic minimum of (L2^2, train index) over the train
 * descriptors sharing a bucket in any of the first nbSchemes schemes (8 in the
 * reference), and symmetrically for train; no candidate -> {~0u, UINT32_MAX}.
 * The max bucket sizes are accepted for interface compatibility and unused. */
int rsfm_cashash_find_best_match(const rsfm_cashash_image* queries, const uint32_t* querySize,
                                 const rsfm_cashash_image* train, uint32_t trainSize,
                                 const rsfm_cashash_result* results, uint32_t nbTasks,
                                 const uint32_t* maxQueryBucketSize, uint32_t maxTrainBucketSize, uint32_t nbSchemes,
                                 hipStream_t stream);

/* Config-1 matching mode (no reference counterpart: the reference computes the
 * 128-bit remap hash, computeHash.cu:474-496, but never matches with it, and its
 * matchWithoutRemap.cu is empty): for every query descriptor the candidates are
 * the train descriptors sharing its bucket in any of the first nbSchemes tables
 * (6 in the cascade-hashing paper); they are ranked by the Hamming distance of
 * the 128-bit remap hashes (ties: lower index), and the first `keep` (8, 16 or
 * 32) are re-ranked by exact L2^2 (ties: lower index). results.train, when
 * set, gets the same search with the roles of the images swapped. No
 * candidate -> {~0u, UINT32_MAX}. At most 2^24 descriptors per image. Device
 * memory: 20 bytes per member and table of every image (the members' indices
 * and remap hashes in bucket order) from the stream-ordered allocator on
 * `stream`, released on it before the call returns. */
typedef struct rsfm_cashash_rerank_image {
    const void* desc;           /* u8 [nbDesc][128], 16-byte aligned */
    const uint8_t* codes;       /* [nbDesc][8] bucket codes, 8-byte aligned */
    const uint32_t* remap;      /* [nbDesc][4] remap hash, 16-byte aligned */
    const uint32_t* bucketEnds; /* [8][256] inclusive prefix */
    const uint32_t* buckets;    /* [8][nbDesc] */
    uint32_t nbDesc;
} rsfm_cashash_rerank_image;
int rsfm_cashash_find_best_match_rerank(const rsfm_cashash_rerank_image* queries,
                                        const rsfm_cashash_rerank_image* train,
                                        const rsfm_cashash_result* results, uint32_t nbTasks, uint32_t nbSchemes,
                                        uint32_t keep, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Match filtering
 * ------------------------------------------------------------------------ */

/* host: crossCheckMatches / removeMatchConflicts (bruteForceMatch.h:126-128,
 * bruteForceMatch.cpp:23-66). out must hold min(nQ, nT) entries; returns the
 * number written. */
size_t rsfm_cross_check_host(const rsfm_best_match* fwd, uint32_t nbQueryDesc,
                             const rsfm_best_match* bwd, uint32_t nbTrainDesc, rsfm_valid_match* out);
size_t rsfm_remove_conflicts_host(const rsfm_best_match* fwd, uint32_t nbQueryDesc,
                                  uint32_t nbTrainDesc, rsfm_valid_match* out);

/* device: keep query i iff fwd[i] valid, (bwd == NULL or bwd[fwd[i].index]
 * .index == i) and (second == NULL or den^2 d1 < num^2 d2, the exact integer
 * ratio test sqrt(d1)/sqrt(d2) < num/den). Output in ascending query order,
 * count written to *count (device pointer). out holds nbQueryDesc entries.
 * workspace: rsfm_filter_matches_workspace_size() bytes. */
int rsfm_filter_matches_workspace_size(uint32_t nbQueryDesc, size_t* bytes);
int rsfm_filter_matches(const rsfm_best_match* fwd, const rsfm_best_match* second,
                        uint32_t nbQueryDesc, const rsfm_best_match* bwd, uint32_t nbTrainDesc,
                        uint32_t ratioNum, uint32_t ratioDen, rsfm_valid_match* out, uint32_t* count,
                        void* workspace, size_t workspaceBytes, hipStream_t stream);

/* one correspondence: (x0, y0) in image 0, (x1, y1) in image 1; the AoS
 * form of the reference's ptPairs (4 x ArrayXf, ImagePair.cpp:148-151) */
typedef struct rsfm_pt_pair {
    float x0, y0, x1, y1;
} rsfm_pt_pair;

/* batched device filter (no reference counterpart; it moves the per-candidate
 * host cross-check of BuilderMethods.cpp:145-202 onto the device): same rule as
 * rsfm_filter_matches per task, one workgroup per task, and optionally gathers
 * the kept keypoint pairs: pts[k] = (kpQuery[q], kpTrain[t]) of the k-th kept
 * match, ready for rsfm_ransac_f8 with nbPtsDev = count. */
typedef struct rsfm_filter_task {
    const rsfm_best_match* fwd;
    const rsfm_best_match* second; /* NULL: no ratio test */
    uint32_t nbQueryDesc;
    const rsfm_best_match* bwd;    /* NULL: no cross-check */
    uint32_t nbTrainDesc;
    rsfm_valid_match* out;         /* optional, nbQueryDesc entries */
    uint32_t* count;               /* device */
    const float* kpQuery;          /* optional (x, y) per query descriptor */
    const float* kpTrain;          /* optional (x, y) per train descriptor */
    rsfm_pt_pair* pts;             /* optional, nbQueryDesc entries */
} rsfm_filter_task;
int rsfm_filter_matches_batch(const rsfm_filter_task* tasks, size_t nbTasks, uint32_t ratioNum, uint32_t ratioDen,
                              hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Affine RANSAC match filter (rsfm::rmf, RansacMatchFilter.cuh:21-34)
 * ------------------------------------------------------------------------ */

/* replaces rmf::cudaRansacMatchFilter (RansacMatchFilter.cu:394-419): matches
 * (x0,y0 = left, x1,y1 = right) reordered by cell with CSR bounds[cols*rows+1];
 * per cell the best of nbRansacTests 3-point affine hypotheses over its 3x3
 * neighbourhood, then (tryOtherAffines) the best of all cells' affines, and a
 * +1 vote for every neighbourhood inlier of a cell whose best affine has >= 6
 * inliers. votes: nbMatches bytes, 4-byte aligned, padded to a multiple of 4.
 * affineScratch: cols*rows*6 floats. seed replaces the reference's clock()
 * seed of the glibc LCG: the thread g of the grid starts at seed + g*22695477. */
int rsfm_rmf_filter(uint8_t* votes, float* affineScratch, const rsfm_pt_pair* matches, uint32_t nbMatches,
                    const uint32_t* bounds, uint32_t cols, uint32_t rows, uint32_t nbRansacTests, float threshold,
                    int tryOtherAffines, uint32_t seed, hipStream_t stream);
/* host: the cell binning of RansacMatchFilter::getInlierMask
 * (RansacMatchFilter.cpp:42-67): cellWidth = ceil(max(w,h)/cellCols), cell =
 * (ceil(y0/cw), ceil(x0/cw)) clamped; reordered/order/bounds are filled
 * (order[k] = original index of reordered[k]); bounds holds cols*rows+1 */
int rsfm_rmf_bin_host(int imgWidth, int imgHeight, uint32_t cellCols, const rsfm_pt_pair* matches, uint32_t nbMatches,
                      uint32_t* cols, uint32_t* rows, uint32_t* cellWidth, rsfm_pt_pair* reordered, uint32_t* bounds,
                      uint32_t* order);

/* Batched form of RansacMatchFilter::getInlierMask + the match selection of its
 * caller (RansacMatchFilter.cpp:40-135, BuilderMethods.cpp:180-193), for pairs
 * whose gathered keypoint pairs are already on the device (e.g. the output of
 * rsfm_filter_matches_batch): per pair, device binning with the reference's
 * rule and order (cellWidth = ceil(max(w, h) / cellCols), cell = clamped ceil
 * of the first point over the cell width), the two cell kernels of
 * rsfm_rmf_filter over all pairs in one launch each (threshold =
 * relativeThreshold * cellWidth, LCG seed per pair), then the pairs with
 * votes >= minVotes are kept: pts (and matches) are compacted in place in
 * their original order and *nbPtsDev becomes the kept count (a pair with fewer
 * than 16 points keeps none, as the reference's filter gives them no votes).
 * Workspace: rsfm_rmf_filter_batch_workspace_size() bytes, 256-byte aligned
 * (per pair: cell-ordered pairs, slots, bounds, affines, votes). */
typedef struct rsfm_rmf_task {
    rsfm_pt_pair* pts;          /* [nbPts] (first image x, y, second image x, y), 16-byte aligned */
    rsfm_valid_match* matches;  /* optional [nbPts], compacted alongside */
    uint32_t* nbPtsDev;         /* device count (capped at nbPts), rewritten with the kept count */
    uint32_t nbPts;             /* capacity */
    uint8_t* inlierMask;        /* optional [nbPts]: votes >= minVotes, original order */
    int imgWidth, imgHeight;    /* first image */
    uint32_t seed;
} rsfm_rmf_task;
int rsfm_rmf_filter_batch_workspace_size(const rsfm_rmf_task* tasks, size_t nbTasks, uint32_t cellCols, size_t* bytes);
int rsfm_rmf_filter_batch(const rsfm_rmf_task* tasks, size_t nbTasks, uint32_t cellCols, uint32_t nbRansacTests,
                          float relativeThreshold, uint32_t minVotes, int tryOtherAffines, void* workspace,
                          size_t workspaceBytes, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Fundamental-matrix RANSAC (8-point hypotheses)
 * ------------------------------------------------------------------------ */

typedef struct rsfm_ransac_task {
    const rsfm_pt_pair* pts;    /* nbPts correspondences, 16-byte aligned */
    uint32_t nbPts;             /* capacity / count when nbPtsDev is NULL */
    const uint32_t* nbPtsDev;   /* optional device count (capped at nbPts) */
    const uint32_t* sampleIdx;  /* [nbHyp][8] indices into pts */
    uint32_t nbHyp;
    uint32_t* counts;           /* [nbHyp] inlier count, 0 for rejected hypotheses */
    float* models;              /* optional [nbHyp][9] row-major F (unit Frobenius) */
    uint32_t* best;             /* optional [2]: best hypothesis (most inliers, lowest index), its count */
    uint8_t* inlierMask;        /* optional [nbPts]: inliers of the best hypothesis */
} rsfm_ransac_task;

/* Replaces the per-hypothesis body of ransac<F,8> in
 * findEpipolarityRansacWithInsufficientTiePts<0> (ImagePair.cpp:148-204;
 * ransac.hpp:115-163) with a batched scorer over given sample indices:
 * sanity gate (squared separations minSepSq0/minSepSq1 in each image, 0 =
 * off; ImagePair.cpp:160-177), FP32 8-point fit (legacy/geometry.hpp:216-259)
 * and inlier count. mode 0: checkEpipolarTransformation's two-sided
 * point-to-line test (legacy/geometry.cpp:55-86) with threshold t pixels;
 * mode 1: Sampson distance < t. */
int rsfm_ransac_f8(const rsfm_ransac_task* tasks, size_t nbTasks, float threshold, int mode, float minSepSq0,
                   float minSepSq1, hipStream_t stream);

/* The end of findEpipolarityRansacWithInsufficientTiePts<0> after the
 * hypotheses of rsfm_ransac_f8 are scored (ImagePair.cpp:185-203, with the
 * adaptive stop of ransac<F,8>, ransac.hpp:139-160):
 *  - selection: with confidence > 0 the scored table is walked in hypothesis
 *    order as the reference's loop walks its draws: zero counts are "insane"
 *    and end the walk after more than min(128, maxIter) in a row; a strictly
 *    larger count becomes the best and sets the required iteration count
 *    min(maxIter, ceil(log2(1 - confidence) / log2(1 - (best / n)^8))); the
 *    walk ends when the sane hypotheses seen reach it, or at the end of the
 *    table. confidence <= 0 takes the whole table. The best hypothesis is the
 *    first with the largest count among the hypotheses walked; with none
 *    (every count 0) the reference's fallback sample {0, ..., 7} is fitted.
 *    The logarithms are the library's own (frexp plus an atanh series in
 *    FP64), so CPU and GPU agree on every cut;
 *  - refit (refit != 0): with >= 8 inliers of the best F, F is refitted by
 *    least squares on all of its inliers (Hartley normalisation and normal
 *    matrix in FP64, null vector by Cholesky and inverse iteration, rank 2 by
 *    a 3x3 Jacobi), and replaces the best F when it has more inliers. The
 *    reference's mask2pts gathers rows by the running count instead of the
 *    point index (legacy/geometry.cpp:341-354), i.e. refits on the first k
 *    points; this entry refits on the inliers, as the call site intends.
 * result: [0] best hypothesis (~0u when every count was 0), [1] inliers of
 * its F, [2] hypotheses walked, [3] inliers of the final F (> [1] exactly
 * when the refit was kept). F is row-major, unit Frobenius (zero when no F
 * could be fitted); inlierMask holds the final F's inliers. n < 8 gives
 * result {~0u, 0, 0, 0}. */
typedef struct rsfm_ransac_finish_task {
    const rsfm_pt_pair* pts;    /* nbPts correspondences, 16-byte aligned */
    uint32_t nbPts;             /* capacity / count when nbPtsDev is NULL */
    const uint32_t* nbPtsDev;   /* optional device count (capped at nbPts) */
    const uint32_t* sampleIdx;  /* [nbHyp][8] the scored samples */
    const uint32_t* counts;     /* [nbHyp] their counts (rsfm_ransac_f8, same threshold and mode) */
    const float* models;        /* optional [nbHyp][9] their models; NULL: the best sample is refitted */
    uint32_t nbHyp;
    float* F;                   /* [9] final F */
    uint32_t* result;           /* [4] */
    uint8_t* inlierMask;        /* optional [nbPts] */
} rsfm_ransac_finish_task;
int rsfm_ransac_f8_finish(const rsfm_ransac_finish_task* tasks, size_t nbTasks, float threshold, int mode,
                          float confidence, uint32_t maxIter, int refit, hipStream_t stream);

/* 4-point homography hypotheses: replaces the per-hypothesis body of
 * ransac<H,4> in findHomographyRansacWithInsufficientTiePts<0>
 * (ImagePair.cpp:349-378): sanity gate (squared separations, as for
 * rsfm_ransac_f8), FP32 4-point DLT fit with Hartley normalisation
 * (findHomography, legacy/geometry.hpp:139-209; null vector by Householder QR)
 * and the inlier count of checkProjectiveTransformation
 * (legacy/geometry.cpp:41-53) with threshold t pixels. sampleIdx is
 * [nbHyp][4]; models are row-major H (x1 ~ H x0), unit Frobenius; best and
 * inlierMask as for rsfm_ransac_f8. */
int rsfm_ransac_h4(const rsfm_ransac_task* tasks, size_t nbTasks, float threshold, float minSepSq0, float minSepSq1,
                   hipStream_t stream);

/* 7-point hypotheses (the sampling model of findEpipolarityDegensac,
 * degensac.cpp:113-150, and findEpipolarityWith7Points,
 * legacy/geometry.cpp:221-307): sampleIdx is [nbHyp][7]; each sample gives up
 * to three F from the cubic singularity constraint (own deterministic solver:
 * FP32 Householder null space, FP64 cubic with bisection roots); a
 * hypothesis's count and model are those of its best F (first on ties).
 * best/inlierMask as for rsfm_ransac_f8. */
int rsfm_ransac_f7(const rsfm_ransac_task* tasks, size_t nbTasks, float threshold, int mode, float minSepSq0,
                   float minSepSq1, hipStream_t stream);

/* DEGENSAC finisher, replacing the per-new-best body of
 * findEpipolarityDegensac (degensac.cpp:126-231) after the 7-point samples
 * were scored by rsfm_ransac_f7 (counts of the same sampleIdx [nbHyp][7]).
 * The nbCandidates best samples (count, then lowest index; zero counts
 * skipped) are replayed in that order: local optimisation of each of the
 * sample's F (checkEpipolarityAndOptimise, degensac.cpp:65-78: FP64
 * least-squares refit on at most 128 inliers evenly spaced by rank, kept when
 * it has more inliers); for a sample beating the best so far, the degeneracy
 * test (H from F and five sample triplets, degensac.cpp:151-187, homography
 * threshold thresholdH); for >= 5 planar sample points, the optimised H
 * (checkHomographyAndOptimise, degensac.cpp:87-100) and a plane-and-parallax
 * search of nbPlaneSamples 2-point F among the points off H
 * (degensac.cpp:195-228; sample s of candidate ordinal c draws two distinct
 * ranks from splitmix64 seeded with seed ^ (c * 4096 + s) * 0xD1B54A32D192ED03).
 * thresholdF and mode are those of rsfm_ransac_f7. Outputs: F and H (row-major,
 * unit Frobenius, zero when none), result = {F inliers, H inliers, 1 when F
 * came from plane and parallax else 0, sample index of the best 7-point F or
 * ~0u}, and the optional inlier masks of F and H (DegensacSolution,
 * degensac.h). At most 65536 correspondences per pair; one workgroup per pair
 * and candidate, whose 132-B records come from hipMallocAsync / hipFreeAsync on
 * the stream. */
typedef struct rsfm_degensac_task {
    const rsfm_pt_pair* pts;    /* nbPts correspondences, 16-byte aligned */
    uint32_t nbPts;             /* capacity / count when nbPtsDev is NULL */
    const uint32_t* nbPtsDev;   /* optional device count (capped at nbPts) */
    const uint32_t* sampleIdx;  /* [nbHyp][7] */
    const uint32_t* counts;     /* [nbHyp] from rsfm_ransac_f7 */
    uint32_t nbHyp;
    float* F;                   /* [9] */
    float* H;                   /* [9] */
    uint32_t* result;           /* [4] */
    uint8_t* inlierMaskF;       /* optional [nbPts] */
    uint8_t* inlierMaskH;       /* optional [nbPts] */
} rsfm_degensac_task;
int rsfm_degensac_refine(const rsfm_degensac_task* tasks, size_t nbTasks, float thresholdF, float thresholdH,
                         int mode, uint32_t nbCandidates, uint32_t nbPlaneSamples, uint64_t seed,
                         hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Relative pose from F (the step after verification)
 * ------------------------------------------------------------------------ */

/* PinHoleCamera (Types.h:231-251): pixel = f * normalised + c */
typedef struct rsfm_camera {
    float fx, fy, cx, cy;
} rsfm_camera;
typedef struct rsfm_pose_task {
    const float* F;             /* [9] row-major, x1^T F x0 = 0 (device) */
    rsfm_camera cam0, cam1;
    const rsfm_pt_pair* pts;    /* nbPts correspondences, 16-byte aligned */
    uint32_t nbPts;
    const uint32_t* nbPtsDev;   /* optional device count (capped at nbPts) */
    const uint8_t* inlierMask;  /* optional [nbPts]: only points with a nonzero mask count */
    float* poses;               /* [4][12]: R row-major then t, x_cam1 = R x_cam0 + t */
    uint32_t* result;           /* [10]: points in front per candidate [4], candidates by
                                   decreasing count [4], number kept, points considered */
} rsfm_pose_task;
/* replaces decomposeEpipolarity (legacy/geometry.hpp:268-346) followed by
 * disambiguateSolutions / isInFrontOfCam (ImagePair.cpp:760-786): E =
 * K1^T F K0 and its four (R, t) in the reference's order (U W^T V^T, U W V^T)
 * x (+U.col(2), -U.col(2)), each scored by the number of correspondences whose
 * triangulated point (legacy/geometry.cpp:146-206) is in front of both
 * cameras; the kept candidates are the leading ones in count order with at
 * least round(n * frontRatio) and at least 2 points in front (the reference's
 * disambiguityFrontRatio, 0.95 by default, Config.h:175). FP64 decomposition,
 * FP32 triangulation; one workgroup per task. */
int rsfm_pose_from_f(const rsfm_pose_task* tasks, size_t nbTasks, float frontRatio, hipStream_t stream);

typedef struct rsfm_pose_h_task {
    const float* H;             /* [9] row-major pixel homography, x1 ~ H x0 (device) */
    rsfm_camera cam0, cam1;
    const rsfm_pt_pair* pts;    /* nbPts correspondences, 16-byte aligned */
    uint32_t nbPts;
    const uint32_t* nbPtsDev;   /* optional device count (capped at nbPts) */
    const uint8_t* inlierMask;  /* optional [nbPts]: only points with a nonzero mask count */
    float* poses;               /* [8][12]: the valid candidates first, R row-major then t (|t| = 1) */
    float* normals;             /* optional [8][3]: plane normal N of each candidate, camera 0 frame */
    uint32_t* result;           /* [19]: points in front per candidate [8], valid candidates by
                                   decreasing count [8], number kept, points considered,
                                   number of valid candidates */
} rsfm_pose_h_task;
/* replaces decomposeHomography (legacy/geometry.hpp:364-447) and the
 * transform to (R, T/d normalised) in solveByHomography /
 * solveByDegensacForBoth (ImagePair.cpp:805-816, 846-856) followed by
 * disambiguateSolutions (ImagePair.cpp:768-786): H normalised by the cameras
 * and its middle singular value, the four (R, N, T/d) of H and the four of -H,
 * those with N . T/d < 1 kept in that order, each scored by the points in
 * front of both cameras, kept = the leading ones with at least
 * round(n * frontRatio) and at least 2. The reference's minNbInliers gate
 * before the decomposition stays with the caller. FP64 decomposition, FP32
 * triangulation; one workgroup per task. */
int rsfm_pose_from_h(const rsfm_pose_h_task* tasks, size_t nbTasks, float frontRatio, hipStream_t stream);

/* RealCamera (Types.h:282-296): pinhole plus distortion {k1, k2, p1, p2, k3} */
typedef struct rsfm_real_camera {
    rsfm_camera pinHole;
    float distortion[5];
} rsfm_real_camera;
/* replaces undistortInPlace (distortion.cpp:251-272): pixel -> normalised
 * coordinates, nbIters Gauss-Newton steps of undistort (distortion.cpp:134-210;
 * the reference uses 2) on the k1, k2, p1, p2, k3 model, back to pixels; px and
 * py (device, nbPts each) are overwritten. Unlike the reference driver, which
 * discards undistort's result (distortion.cpp:238), the undistorted points are
 * written. All-zero distortion leaves the points untouched. */
int rsfm_undistort_in_place(float* px, float* py, size_t nbPts, const rsfm_real_camera* camera, uint32_t nbIters,
                            hipStream_t stream);

/* the same undistortion for gathered keypoint pairs, as makePtPair does it
 * (ImagePair.cpp:105-112: the first point of each pair with the first image's
 * camera, the second with the second's); a side whose distortion is all zero
 * is left untouched; many pairs per launch */
typedef struct rsfm_undistort_task {
    rsfm_pt_pair* pts;          /* nbPts keypoint pairs, overwritten, 16-byte aligned (device) */
    uint32_t nbPts;
    const uint32_t* nbPtsDev;   /* optional device count (capped at nbPts) */
    rsfm_real_camera cam0, cam1;
} rsfm_undistort_task;
int rsfm_undistort_pt_pairs(const rsfm_undistort_task* tasks, size_t nbTasks, uint32_t nbIters, hipStream_t stream);

/* deterministic device sampler (no reference counterpart: the reference's
 * engine is random_device-seeded, ransac.hpp:117): hypothesis h draws 8
 * distinct indices from splitmix64 seeded with seed ^ h * 0xD1B54A32D192ED03,
 * rejecting repeats; fewer than 8 points give ~0u entries */
int rsfm_ransac_sample_indices(uint32_t* sampleIdx, uint32_t nbHyp, uint32_t nbPts, const uint32_t* nbPtsDev,
                               uint64_t seed, hipStream_t stream);

/* the same sampler for many pairs in one launch */
typedef struct rsfm_sample_task {
    uint32_t* sampleIdx; /* [nbHyp][8] */
    uint32_t nbHyp;
    uint32_t nbPts;
    const uint32_t* nbPtsDev; /* optional device count (capped at nbPts) */
    uint64_t seed;
} rsfm_sample_task;
int rsfm_ransac_sample_indices_batch(const rsfm_sample_task* tasks, size_t nbTasks, hipStream_t stream);
/* the same draw with sampleSize (4..8) indices per hypothesis: sample h holds
 * the first sampleSize distinct draws of hypothesis h's sequence, rows of
 * sampleSize (7 for rsfm_ransac_f7, 4 for rsfm_ransac_h4); fewer points than
 * sampleSize give ~0u entries */
int rsfm_ransac_sample_indices_batch_k(const rsfm_sample_task* tasks, size_t nbTasks, uint32_t sampleSize,
                                       hipStream_t stream);

/* ---- batched pipeline (no reference counterpart: the reference matches each
 * candidate pair with BruteForceMatcher::match, bruteForceMatcher.cpp:163-443,
 * and verifies it on a host fiber, BuilderMethods.cpp:141-202 /
 * ImagePair.cpp:148-204). Many pairs of uneven size are matched (top-2 query
 * side, top-1 train side), filtered (ratio test + cross-check + keypoint
 * gather), sampled and scored by one HIP graph launch. All buffers are the
 * caller's except the pipeline's own task tables and partials. Inputs must be
 * complete when rsfm_pipeline_create is called: it times candidate matcher
 * tilings on them (autotuning) on a private stream before capturing the graph.
 * Results are the same as the individual entry points' (rsfm_bf_match_top2,
 * rsfm_filter_matches_batch, rsfm_ransac_sample_indices_batch with seed + pair
 * index, rsfm_ransac_f8 with nbPtsDev = count). ---- */
typedef struct rsfm_pipeline_pair {
    uint32_t nA, nB;
    const void* descA;          /* device u8 [nA][128], 16-byte aligned */
    const void* descB;          /* device u8 [nB][128], 16-byte aligned */
    const float* kpA;           /* device (x, y) per query keypoint */
    const float* kpB;           /* device (x, y) per train keypoint */
    rsfm_best_match* bestA;     /* [nA] */
    rsfm_best_match* secondA;   /* [nA] */
    rsfm_best_match* bestB;     /* [nB] */
    rsfm_valid_match* matches;  /* optional [nA] kept matches */
    uint32_t* count;            /* [1] kept matches */
    rsfm_pt_pair* pts;          /* [nA] kept keypoint pairs, 16-byte aligned */
    uint32_t* sampleIdx;        /* [nbHyp][8] */
    uint32_t* counts;           /* [nbHyp] inlier counts */
    uint32_t* best;             /* [2] best hypothesis, its count */
    uint8_t* inlierMask;        /* optional [nA]; with the DEGENSAC solver F's inliers */
    /* DEGENSAC solver only (cfg.solver = RSFM_PIPELINE_SOLVER_DEGENSAC): sampleIdx is then
     * [nbHyp][7], counts are rsfm_ransac_f7's and best its best sample; the finisher
     * (rsfm_degensac_refine) writes these */
    float* F;                   /* [9] */
    float* H;                   /* [9] */
    uint32_t* solverResult;     /* [4]: F inliers, H inliers, plane-and-parallax flag, sample */
    uint8_t* inlierMaskH;       /* optional [nA] */
    /* optional pose stage, DEGENSAC solver only: rsfm_pose_from_f over F's inliers when
     * posesF is set, rsfm_pose_from_h over H's inliers when posesH is set */
    rsfm_camera cam0, cam1;
    float* posesF;              /* [4][12] */
    uint32_t* poseResultF;      /* [10] */
    float* posesH;              /* [8][12] */
    float* normalsH;            /* optional [8][3] */
    uint32_t* poseResultH;      /* [19] */
    /* optional host [10]: k1, k2, p1, p2, k3 of camera 0, then of camera 1 (read at
     * creation); when set, the gathered keypoint pairs are undistorted in place with
     * cam0/cam1 (rsfm_undistort_pt_pairs, 2 Gauss-Newton steps as the reference's
     * makePtPair) before sampling */
    const float* distortion;
    /* first image size, for the optional affine match filter stage */
    int width0, height0;
} rsfm_pipeline_pair;
typedef struct rsfm_pipeline_config {
    uint32_t nbHyp;            /* hypotheses per pair */
    float threshold;           /* inlier threshold in pixels */
    int mode;                  /* 0 two-sided line test, 1 Sampson */
    uint32_t ratioNum, ratioDen; /* ratio test num/den (4, 5 = 0.8) */
    uint64_t seed;             /* pair p samples with seed + p */
    float minSepSq0, minSepSq1;  /* sample sanity gate, 0 = off */
    uint32_t segCols;          /* matcher tile width (multiple of 64, <= 4096); 0 = autotune */
    uint32_t wgPerXcd;         /* persistent matcher workgroups per XCD, RSFM_PIPELINE_WG_PER_TILE
                                  (one workgroup per tile, no persistent loop); 0 = autotune */
    uint32_t solver;           /* RSFM_PIPELINE_SOLVER_F8 or RSFM_PIPELINE_SOLVER_DEGENSAC */
    float thresholdH;          /* DEGENSAC: homography threshold in pixels */
    uint32_t nbCandidates;     /* DEGENSAC: best 7-point samples replayed */
    uint32_t nbPlaneSamples;   /* DEGENSAC: plane-and-parallax samples per candidate */
    float frontRatio;          /* pose stage: disambiguation ratio (the reference's 0.95) */
    /* 8-point solver: the end of the reference's driver (rsfm_ransac_f8_finish) as a graph
     * stage when refit != 0 or confidence > 0; every pair then needs F and solverResult,
     * which receive the final F and the finisher's result [4], and inlierMask holds the
     * final F's inliers */
    float confidence;          /* adaptive cut of the hypothesis table, <= 0 = whole table */
    uint32_t maxIter;          /* the reference's max_iter (5000) for the cut */
    uint32_t refit;            /* refit F on the best sample's inliers */
    /* optional affine match filter stage (rsfm_rmf_filter_batch, the reference's default
     * RansacMatchFilter before pair solving, BuilderMethods.cpp:180-193) between the
     * ratio/cross-check filter and the solver: the kept matches and keypoint pairs are
     * compacted in place and count becomes the filtered count; LCG seed = seed + pair */
    uint32_t matchFilter;      /* 0 off, 1 on (every pair then needs width0, height0) */
    uint32_t mfCellCols;       /* 16 */
    uint32_t mfTests;          /* 256 */
    float mfThreshold;         /* relative to the cell width, 0.15 */
    uint32_t mfMinVotes;       /* 2 */
    uint32_t mfTryOther;       /* 1 */
} rsfm_pipeline_config;
#define RSFM_PIPELINE_WG_PER_TILE 0xFFFFFFFFu
/* solvers: 8-point F RANSAC (samples, scores, best and inlier mask), or the
 * reference's default pair solver (Config.h:169, solveByDegensacForBoth,
 * ImagePair.cpp:833-866): 7-point samples drawn by
 * rsfm_ransac_sample_indices_batch_k, rsfm_ransac_f7, rsfm_degensac_refine
 * (seed = cfg.seed) and the optional pose stage, with the pair's match count
 * as the device point count */
#define RSFM_PIPELINE_SOLVER_F8 0u
#define RSFM_PIPELINE_SOLVER_DEGENSAC 1u
typedef struct rsfm_pipeline_tuning_entry {
    uint32_t segCols, wgPerXcd;
    float matchMs;             /* fill + match + merge, best of three */
} rsfm_pipeline_tuning_entry;
typedef struct rsfm_pipeline_info_t {
    uint32_t segCols, wgPerXcd, nbTiles;
    float tunedMatchMs;        /* 0 when nothing was tuned */
} rsfm_pipeline_info_t;
typedef struct rsfm_pipeline rsfm_pipeline;
int rsfm_pipeline_create(const rsfm_pipeline_pair* pairs, size_t nbPairs, const rsfm_pipeline_config* cfg,
                         rsfm_pipeline** out);
/* one run in flight at a time: the graph's stages share the pipeline's task tables and
 * partials; rsfm_pipeline_destroy waits for the last run's stream */
int rsfm_pipeline_run(rsfm_pipeline* pipeline, hipStream_t stream);
int rsfm_pipeline_info(const rsfm_pipeline* pipeline, rsfm_pipeline_info_t* info, rsfm_pipeline_tuning_entry* entries,
                       size_t maxEntries, size_t* nbEntries);
int rsfm_pipeline_destroy(rsfm_pipeline* pipeline);

/* ---- learned proposals (no reference counterpart: the reference samples
 * uniformly, NonDuplicateSampler in ransac.hpp:23-82) ---- */

/* weighted sampler: hypothesis h draws 8 distinct indices, each draw with
 * probability proportional to q_i = round(clamp(weights[i], 0, 1) * 65536)
 * among the points not drawn yet (same splitmix64 stream as
 * rsfm_ransac_sample_indices; 512 draws at most, then ~0u). All-zero weights
 * fall back to the uniform draw. cdf: device workspace of nbPts u64. */
typedef struct rsfm_weighted_sample_task {
    uint32_t* sampleIdx; /* [nbHyp][8] */
    uint32_t nbHyp;
    uint32_t nbPts;
    const uint32_t* nbPtsDev; /* optional device count (capped at nbPts) */
    const float* weights;     /* [nbPts] */
    uint64_t* cdf;            /* [nbPts] workspace */
    uint64_t seed;
} rsfm_weighted_sample_task;
int rsfm_ransac_sample_weighted_batch(const rsfm_weighted_sample_task* tasks, size_t nbTasks, hipStream_t stream);

/* input features of the proposal network for each kept match k < *count:
 * f0 = sqrt(d1 / d2) of the ratio test (0 without a second match, 1 if d2 =
 * 0), f1 = sqrt(d1) / 512, f2..f5 = query and train keypoints mapped to
 * [-1, 1] (x / width * 2 - 1, y / height * 2 - 1), f6 = log2(train scale /
 * query scale) clamped to [-8, 8], f7 = orientation difference wrapped to
 * [-pi, pi) over pi. Keypoints are (x, y, scale, orientation) float4. Rows
 * count .. maxMatches-1 are zero. */
typedef struct rsfm_feature_task {
    const rsfm_valid_match* matches; /* from rsfm_filter_matches_batch */
    const uint32_t* count;           /* device */
    uint32_t maxMatches;
    const rsfm_best_match* second;   /* second best per query */
    const float* kpQuery;            /* [nQuery][4] */
    const float* kpTrain;            /* [nTrain][4] */
    float width, height;
    float* x;                        /* [maxMatches][8] */
} rsfm_feature_task;
int rsfm_proposal_features_batch(const rsfm_feature_task* tasks, size_t nbTasks, hipStream_t stream);

/* proposal network 8 -> 64 -> 64 -> 64 -> 1 (ReLU, sigmoid) on bf16 matrix
 * cores with FP32 accumulation. params: FP32 vector of rsfm_mlp_param_count()
 * entries, W1[64x8] b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[64] b4,
 * row-major W[out][in]; weights and inputs are rounded to bf16 on load.
 * x: [n][8] FP32 features, 16-byte aligned. */
int rsfm_mlp_param_count(void);
int rsfm_mlp_forward(const float* params, const float* x, uint32_t n, float* prob, hipStream_t stream);
/* gradient of the mean binary cross-entropy over the n samples (labels y in
 * {0, 1}) into grad (same layout as params) and the mean loss into *loss
 * (optional, device). Deterministic: per-workgroup partials summed in order. */
int rsfm_mlp_grad_workspace_size(uint32_t n, size_t* bytes);
int rsfm_mlp_grad(const float* params, const float* x, const float* y, uint32_t n, float* grad, float* loss,
                  void* workspace, size_t workspaceBytes, hipStream_t stream);
/* one Adam step (step counts from 1) on the FP32 master parameters */
int rsfm_mlp_adam_step(float* params, const float* grad, float* m, float* v, uint32_t step, float lr, float beta1,
                       float beta2, float eps, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
