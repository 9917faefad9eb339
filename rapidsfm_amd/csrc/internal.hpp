// internal.hpp - task descriptors shared between the kernels of the separate
// translation units and the batched pipeline (library-internal).
#pragma once
#include "common.hpp"

namespace rsfm_amd {

// task source of a kernel: a by-value batch (kernel argument) or a table in
// device memory (batched pipeline); kernels index either with src[i]
template <class T, int kMax>
struct ByValue {
    T t[kMax];
    __device__ const T& operator[](uint32_t i) const { return t[i]; }
};
template <class T>
struct InTable {
    const T* p;
    __device__ const T& operator[](uint32_t i) const { return p[i]; }
};

namespace bf {
constexpr int kBM = 256;                // query rows per workgroup tile
constexpr int kBN = 64;                 // train columns per staged chunk
constexpr uint32_t kMaxSegColsTop2 = 4096;  // columns per segment: 256 blocks of 16, the 8-bit tag of a row key
constexpr int kMaxTasks = 32;

struct DevTask {
    uint32_t nA, nB;
    const uint8_t* A;
    const uint8_t* B;
    uint64_t* bestA;
    uint64_t* secondA;  // top-2 only
    uint64_t* bestB;    // may be null
    uint64_t* part;     // top-2 with several segments: [nSeg][nA][2]
    uint32_t nSeg;      // column segments of this task (1: results written directly)
    uint32_t segCols;   // columns per segment (multiple of kBN, <= kMaxSegColsTop2)
    // staging form of B, written by the preparation kernel: w = b - 128
    // as signed bytes for nB rounded up to kBN columns (zero past nB), per
    // column NB' << 8 with NB' = |b - 127|^2 - 128 (a fixed offset past nB)
    // and |b|^2
    uint8_t* W;
    int32_t* NB;
    // top-2 train side: per train column the smallest d - p over the query
    // row blocks (merged by atomicMin)
    int32_t* colMin;
};
typedef ByValue<DevTask, kMaxTasks> TaskBatch;
typedef InTable<DevTask> TaskTable;

// one workgroup tile of the batched pipeline: kBM query rows x one segment
struct TileDesc {
    uint32_t task, rowBase, colBeg, seg;
};

// persistent top-2 matcher over a tile table split into kNumXcd lists
int launch_match_tiles_top2(const DevTask* tasks, const TileDesc* tiles, const uint32_t* xcdOff, uint32_t grid,
                            hipStream_t stream);
// one workgroup per tile of the same table; maxList = longest XCD list
int launch_match_tile_grid_top2(const DevTask* tasks, const TileDesc* tiles, const uint32_t* xcdOff, uint32_t maxList,
                                hipStream_t stream);
// workspace bytes of one top-2 task's staging form and train-side partials
// (W, NB, colMin; 256-B aligned parts)
size_t prep_bytes(uint32_t nA, uint32_t nB);
// point W, NB and colMin of t into ws (prep_bytes(t.nA, t.nB) bytes)
void place_prep(DevTask& t, uint8_t* ws);
// staging form of B for the listed tasks (W, NB set, colMin cleared); maxB = largest nB
int launch_prep_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxB, hipStream_t stream);

// merge of the segment partials of the listed tasks (all with nSeg > 1)
int launch_merge_top2_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxA, hipStream_t stream);
}  // namespace bf

namespace aux {
struct FilterTask {
    const uint64_t* fwd;
    const uint64_t* second;
    uint32_t nQ;
    const uint64_t* bwd;
    uint32_t nT;
    rsfm_valid_match* out;
    uint32_t* count;
    const float2* kpQ;
    const float2* kpT;
    float4* pts;
};
int launch_filter_table(const FilterTask* tasks, uint32_t nTasks, uint32_t num, uint32_t den, hipStream_t stream);
}  // namespace aux

namespace ransac {
struct DevTask {
    const float4* pts;
    const uint32_t* nDev;
    uint32_t nHost;
    uint32_t H;
    const uint32_t* idx;
    uint32_t* counts;
    float* models;
    uint32_t* best;
    uint8_t* mask;
};
struct SampleTask {
    uint32_t* idx;
    uint32_t H;
    uint32_t nHost;
    const uint32_t* nDev;
    uint64_t seed;
    uint32_t k;  // indices per sample (4..8), the first k draws of the sequence
};
int launch_sample_table(const SampleTask* tasks, uint32_t nTasks, uint32_t maxH, hipStream_t stream);
// scores (and models) of every hypothesis; withBest also launches the best-model / mask kernel
int launch_score_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxH, float threshold, int mode,
                       float minSepSq0, float minSepSq1, hipStream_t stream, bool withBest = true);
// end of the 8-point driver (rsfm_ransac_f8_finish): adaptive cut, refit on the inliers
struct FinishTask {
    const float4* pts;
    const uint32_t* nDev;
    uint32_t nHost;
    const uint32_t* idx;
    const uint32_t* counts;
    const float* models;
    uint32_t H;
    float* F;
    uint32_t* result;
    uint8_t* mask;
    uint32_t* best;  // optional [2]: best hypothesis, its count
};
struct FinishParams {
    float t2;
    float conf;  // <= 0: the whole table
    uint32_t maxIter;
    int refit;
};
int make_finish_params(float threshold, int mode, float confidence, uint32_t maxIter, int refit, FinishParams& p);
int launch_finish_table(const FinishTask* tasks, uint32_t nTasks, const FinishParams& p, int mode, hipStream_t stream);
// 7-point scoring + best over a device task table (sampleIdx rows of 7)
int launch_score7_table(const DevTask* tasks, uint32_t nTasks, uint32_t maxH, float threshold, int mode,
                        float minSepSq0, float minSepSq1, hipStream_t stream);
}  // namespace ransac

namespace rmf {
// batched affine match filter (rsfm_rmf_filter_batch) over one pair
struct BatchTask {
    float4* pts;                // gathered pairs, compacted in place
    rsfm_valid_match* matches;  // optional, compacted alongside
    uint32_t* count;            // device count, rewritten with the kept count
    uint32_t nHost;
    uint8_t* mask;              // optional, original order
    float4* reordered;          // workspace: pairs in cell order
    uint32_t* pos;              // workspace: slot of point i in cell order
    uint32_t* bounds;           // workspace: [cells + 1]
    float* affine;              // workspace: [cells][6]
    uint8_t* votes;             // workspace: [nHost rounded up to 4], cell order
    uint32_t cols, rows, cellWidth;
    float threshold;            // pixels
    uint32_t seed;
};
struct BatchParams {
    uint32_t nbTests, minVotes;
    int tryOther;
};
// argument checks and grid of one pair; its workspace bytes; its workspace pointers
int make_batch_task(const rsfm_rmf_task& t, uint32_t cellCols, float relativeThreshold, BatchTask& out);
size_t batch_workspace_bytes(uint32_t nPts, uint32_t nCells);
void place_workspace(BatchTask& t, uint8_t*& ws);
// binning, cell kernels and compaction over a device task table
int launch_batch_table(const BatchTask* tasks, uint32_t nTasks, uint32_t maxCols, uint32_t maxRows, const BatchParams& p,
                       hipStream_t stream);
}  // namespace rmf

namespace undist {
struct Params {
    float fx, fy, cx, cy;      // pinhole
    float ifx, ify, icx, icy;  // inverse pinhole: 1/f, -c/f
    float k1, k2, p1, p2, k3;
    uint32_t iters;
};
struct PairTask {
    float4* pts;
    const uint32_t* nDev;
    uint32_t nHost;
    int on0, on1;  // side has nonzero distortion
    Params d0, d1;
};
// argument checks of rsfm_undistort_in_place; active = any nonzero distortion parameter
int make_params(const rsfm_real_camera& camera, uint32_t nbIters, Params& d, bool& active);
// one launch over a device table of keypoint-pair tasks
int launch_pairs_table(const PairTask* tasks, uint32_t nTasks, uint32_t maxN, hipStream_t stream);
}  // namespace undist

namespace pose {
struct DevTask {
    const float* M;          // F or H
    float cam0[4], cam1[4];  // fx, fy, cx, cy
    const float4* pts;
    const uint32_t* nDev;
    uint32_t nHost;
    const uint8_t* mask;
    float* poses;
    float* normals;  // H only, optional
    uint32_t* result;
};
// argument checks of rsfm_pose_from_f / rsfm_pose_from_h for one task
int make_task(const rsfm_pose_task& t, DevTask& d);
int make_task(const rsfm_pose_h_task& t, DevTask& d);
// one workgroup per task of a device table; kind 0 = F, 1 = H
int check_ratio(float frontRatio);
int launch_pose_table(const DevTask* tasks, uint32_t nTasks, int kind, float frontRatio, hipStream_t stream);
}  // namespace pose

namespace degensac {
// what the per-new-best body of the reference computes for one candidate
// sample; it depends on the candidate alone, so the candidates are worked in
// parallel and their records replayed in order afterwards
struct CandRec {
    uint64_t planeKey;  // best plane-and-parallax sample: count << 32 | ~index (0: none usable)
    uint32_t h;       // sample (hypothesis) index; ~0u: no candidate
    uint32_t cntF;    // optimised count of the sample's best F (0: none usable)
    uint32_t planar;  // the degeneracy test put >= 5 of the 7 points on one H
    uint32_t cntH;    // optimised count of that H
    uint32_t plane;   // the plane-and-parallax search gave an F
    uint32_t cntFh;   // its optimised count
    float F[9], H[9], Fh[9];
};
struct DevTask {
    const float4* pts;
    const uint32_t* nDev;
    uint32_t nHost;
    const uint32_t* idx;
    const uint32_t* counts;
    uint32_t H;
    float* F;
    float* Hm;
    uint32_t* result;
    uint8_t* maskF;
    uint8_t* maskH;
    CandRec* cand;  // [nCand] workspace
};
struct Params {
    float tF2, tH2;
    uint32_t nCand, nPlane;
    uint64_t seed;
};
// argument checks of rsfm_degensac_refine: the shared parameters, then one task
int make_params(float thresholdF, float thresholdH, int mode, uint32_t nbCandidates, uint32_t nbPlaneSamples,
                uint64_t seed, Params& out);
int make_task(const rsfm_degensac_task& t, DevTask& out);
// one workgroup per task of a device task table, all tasks in one launch
int launch_degensac_table(const DevTask* tasks, uint32_t nTasks, const Params& p, int mode, hipStream_t stream);
}  // namespace degensac

}  // namespace rsfm_amd
