// pipeline.hip - batched two-view pipeline: many image pairs of uneven size
// matched and verified by one HIP graph launch.
//
// No reference counterpart: the reference runs each candidate pair through
// BruteForceMatcher::match (bruteForceMatcher.cpp:163-443, <=16 queries per
// launch and a stream sync per launch group) and then verifies it on a host
// fiber (BuilderMethods.cpp:141-202, ImagePair.cpp:148-204).
//
// Design (own, for MI355X):
//  * All per-pair descriptors live in device tables (bf::DevTask,
//    aux::FilterTask, ransac::SampleTask / DevTask) uploaded once at creation,
//    so every stage covers the whole batch in one kernel launch regardless of
//    the kernel-argument limit.
//  * The matcher runs persistently over a tile table: a tile is 256 query rows
//    x one column segment of one pair. Pairs are assigned to the 8 XCDs by
//    longest-processing-time first (work nA x nB), so every XCD gets the same
//    amount of work and all tiles of a pair share one L2; inside an XCD the
//    pairs run largest first. Tile width (segment columns) and the persistent
//    grid are chosen by an autotuner that times the candidates on the batch
//    itself (HIP events, best of three) unless the caller fixes them.
//  * The stages (fill, match, segment merge, ratio/cross-check filter with
//    keypoint gather, seeded sampler, 8-point fit + scoring, best model and
//    inlier mask) are captured once into a HIP graph on a private stream;
//    rsfm_pipeline_run launches that graph. Buffers are cleared by a fill
//    kernel (no memset nodes).
//  * Pairs that carry lens distortion have their gathered keypoint pairs
//    undistorted in place right after the filter (one launch over a device
//    table), as the reference's makePtPair does before solving.
//  * An optional affine match filter stage (the reference's RansacMatchFilter
//    before pair solving) bins, filters and compacts every pair's gathered
//    keypoint pairs on the device (rmf.hip), four launches for the batch.
//  * With the 8-point solver the end of the reference's driver can follow
//    the scorer as one more stage (adaptive cut, refit on the inliers;
//    ransac_finish.hip), one workgroup per pair.
//  * With the DEGENSAC solver the verification stages are the reference's
//    default pair solver instead: 7-point samples, the 7-point scorer, the
//    DEGENSAC finisher (F and H with their inlier masks) and, optionally, the
//    relative pose from F and from H, captured into the same graph. The
//    7-point scorer, the finisher and the pose stages read device task
//    tables, so each covers every pair in one launch (one finisher or pose
//    workgroup per pair).
#include "internal.hpp"
#include <algorithm>
#include <numeric>
#include <vector>

namespace rsfm_amd {
namespace pipe {

struct FillRange {
    uint64_t* p;
    uint32_t n;
};

__global__ void fill_kernel(const FillRange* __restrict__ ranges, uint64_t value)
{
    const FillRange r = ranges[blockIdx.y];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < r.n; i += gridDim.x * blockDim.x) r.p[i] = value;
}

}  // namespace pipe
}  // namespace rsfm_amd

using namespace rsfm_amd;

struct rsfm_pipeline {
    std::vector<rsfm_pipeline_pair> pairs;
    rsfm_pipeline_config cfg{};
    hipStream_t stream = nullptr;  // capture stream
    uint8_t* arena = nullptr;
    // device tables
    bf::DevTask* dTasks = nullptr;
    bf::DevTask* dMerge = nullptr;
    uint32_t nMerge = 0, maxAMerge = 0;
    bf::TileDesc* dTiles = nullptr;
    uint32_t* dXcdOff = nullptr;
    size_t tileCap = 0;
    aux::FilterTask* dFilter = nullptr;
    ransac::SampleTask* dSample = nullptr;
    ransac::DevTask* dRansac = nullptr;
    ransac::FinishTask* dFinish = nullptr;  // 8-point solver with the finishing stage
    ransac::FinishParams finParams{};
    bool finish = false;
    rmf::BatchTask* dRmf = nullptr;  // affine match filter stage
    rmf::BatchParams rmfParams{};
    uint32_t rmfCols = 0, rmfRows = 0;
    pipe::FillRange* dFill = nullptr;
    uint32_t nFill = 0, maxFill = 0;
    uint64_t* dPart = nullptr;
    size_t partCap = 0;
    uint8_t* dPrep = nullptr;  // staging form of every pair's train descriptors
    uint32_t maxPrepB = 0;
    // chosen tiling
    uint32_t segCols = 0, wgPerXcd = 0, nTiles = 0, maxList = 0;
    float tunedMs = 0.f;
    std::vector<rsfm_pipeline_tuning_entry> tuning;
    // DEGENSAC solver stages: device task tables for the 7-point scorer and the
    // finisher (every pair in one launch), host task arrays for the pose entries
    ransac::DevTask* dF7 = nullptr;
    degensac::DevTask* dDg = nullptr;
    degensac::Params dgParams{};
    undist::PairTask* dUnd = nullptr;  // pairs with distortion
    uint32_t nUnd = 0, maxUnd = 0;
    pose::DevTask* dPoseF = nullptr;
    pose::DevTask* dPoseH = nullptr;
    uint32_t nPoseF = 0, nPoseH = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;  // recorded on the stream of every run
    bool ran = false;
};

namespace {

uint32_t tasks_max_a(const std::vector<rsfm_pipeline_pair>& p)
{
    uint32_t m = 0;
    for (auto& x : p) m = std::max(m, x.nA);
    return m;
}

// Host plan for one segment width: per-pair task descriptors (segments, part
// offsets), the tile table in XCD lists and the total partial-buffer size.
struct Plan {
    std::vector<bf::DevTask> tasks;
    std::vector<bf::TileDesc> tiles;
    std::vector<uint32_t> xcdOff;
    size_t partWords = 0;
    size_t prepBytes = 0;
};

Plan make_plan(const std::vector<rsfm_pipeline_pair>& pairs, uint32_t segCols, uint64_t* part, uint8_t* prep)
{
    Plan pl;
    const size_t P = pairs.size();
    pl.tasks.resize(P);
    for (size_t i = 0; i < P; i++) {
        const auto& x = pairs[i];
        bf::DevTask& t = pl.tasks[i];
        t = bf::DevTask{};
        t.nA = x.nA;
        t.nB = x.nB;
        t.A = static_cast<const uint8_t*>(x.descA);
        t.B = static_cast<const uint8_t*>(x.descB);
        t.bestA = reinterpret_cast<uint64_t*>(x.bestA);
        t.secondA = reinterpret_cast<uint64_t*>(x.secondA);
        t.bestB = reinterpret_cast<uint64_t*>(x.bestB);
        t.segCols = segCols;
        t.nSeg = std::max(1u, div_up(x.nB, segCols));
        if (t.nSeg > 1) {
            t.part = part ? part + pl.partWords : nullptr;
            pl.partWords += (size_t)t.nSeg * x.nA * 2;
        }
        if (prep) bf::place_prep(t, prep + pl.prepBytes);
        pl.prepBytes += bf::prep_bytes(x.nA, x.nB);
    }
    // pairs to XCDs: longest processing time first
    std::vector<size_t> order(P);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return (uint64_t)pairs[a].nA * pairs[a].nB > (uint64_t)pairs[b].nA * pairs[b].nB;
    });
    std::vector<uint64_t> load(kNumXcd, 0);
    std::vector<std::vector<size_t>> byXcd(kNumXcd);
    for (size_t i : order) {
        const size_t x = std::min_element(load.begin(), load.end()) - load.begin();
        load[x] += (uint64_t)pairs[i].nA * pairs[i].nB + 1;
        byXcd[x].push_back(i);
    }
    pl.xcdOff.assign(kNumXcd + 1, 0);
    for (uint32_t x = 0; x < kNumXcd; x++) {
        pl.xcdOff[x] = (uint32_t)pl.tiles.size();
        for (size_t i : byXcd[x]) {
            const bf::DevTask& t = pl.tasks[i];
            if (t.nA == 0 || t.nB == 0) continue;
            for (uint32_t s = 0; s < t.nSeg; s++)
                for (uint32_t r = 0; r < div_up(t.nA, bf::kBM); r++)
                    pl.tiles.push_back(bf::TileDesc{(uint32_t)i, r * bf::kBM, s * segCols, s});
        }
    }
    pl.xcdOff[kNumXcd] = (uint32_t)pl.tiles.size();
    return pl;
}

int upload(void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return RSFM_OK;
    return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "pipeline upload");
}

// fill + match + merge of the current plan on stream s
int enqueue_match(rsfm_pipeline* pl, hipStream_t s)
{
    if (pl->nFill) {
        pipe::fill_kernel<<<dim3(std::min<uint32_t>(div_up(pl->maxFill, 256), 64), pl->nFill), 256, 0, s>>>(pl->dFill,
                                                                                                         kNoMatch);
        RSFM_CHECK(check_launch("pipeline fill"));
    }
    if (pl->nTiles) {
        RSFM_CHECK(bf::launch_prep_table(pl->dTasks, (uint32_t)pl->pairs.size(), pl->maxPrepB, s));
        if (pl->wgPerXcd == RSFM_PIPELINE_WG_PER_TILE)
            RSFM_CHECK(bf::launch_match_tile_grid_top2(pl->dTasks, pl->dTiles, pl->dXcdOff, pl->maxList, s));
        else
            RSFM_CHECK(bf::launch_match_tiles_top2(pl->dTasks, pl->dTiles, pl->dXcdOff, pl->wgPerXcd * kNumXcd, s));
    }
    RSFM_CHECK(bf::launch_merge_top2_table(pl->dMerge, pl->nMerge, pl->maxAMerge, s));
    return RSFM_OK;
}

// upload the plan for segment width segCols
int apply_plan(rsfm_pipeline* pl, uint32_t segCols)
{
    Plan p = make_plan(pl->pairs, segCols, pl->dPart, pl->dPrep);
    RSFM_RETURN_IF(p.partWords > pl->partCap, RSFM_ERR_INVALID, "pipeline: partial buffer too small");
    RSFM_RETURN_IF(p.tiles.size() > pl->tileCap, RSFM_ERR_INVALID, "pipeline: tile table too small");
    RSFM_CHECK(upload(pl->dTasks, p.tasks.data(), p.tasks.size() * sizeof(bf::DevTask)));
    RSFM_CHECK(upload(pl->dTiles, p.tiles.data(), p.tiles.size() * sizeof(bf::TileDesc)));
    RSFM_CHECK(upload(pl->dXcdOff, p.xcdOff.data(), p.xcdOff.size() * sizeof(uint32_t)));
    std::vector<bf::DevTask> merge;
    uint32_t maxA = 0;
    for (auto& t : p.tasks)
        if (t.nSeg > 1 && t.nA > 0) {
            merge.push_back(t);
            maxA = std::max(maxA, t.nA);
        }
    RSFM_CHECK(upload(pl->dMerge, merge.data(), merge.size() * sizeof(bf::DevTask)));
    pl->nMerge = (uint32_t)merge.size();
    pl->maxAMerge = maxA;
    pl->segCols = segCols;
    pl->nTiles = (uint32_t)p.tiles.size();
    pl->maxList = 0;
    for (uint32_t x = 0; x < kNumXcd; x++) pl->maxList = std::max(pl->maxList, p.xcdOff[x + 1] - p.xcdOff[x]);
    return RSFM_OK;
}

int tune(rsfm_pipeline* pl)
{
    std::vector<uint32_t> segs, wgs;
    if (pl->cfg.segCols)
        segs.push_back(pl->cfg.segCols);
    else
        segs = {1024, 2048, 4096};
    if (pl->cfg.wgPerXcd)
        wgs.push_back(pl->cfg.wgPerXcd);
    else
        wgs = {32, 64, RSFM_PIPELINE_WG_PER_TILE};
    const bool timed = segs.size() * wgs.size() > 1;
    hipEvent_t e0, e1;
    if (timed) {
        RSFM_CHECK(check_hip(hipEventCreate(&e0), "event"));
        RSFM_CHECK(check_hip(hipEventCreate(&e1), "event"));
    }
    float bestMs = 1e30f;
    uint32_t bestSeg = segs[0], bestWg = wgs[0];
    int rc = RSFM_OK;
    for (uint32_t sc : segs) {
        if (rc != RSFM_OK) break;
        if ((rc = apply_plan(pl, sc)) != RSFM_OK) break;
        for (uint32_t wg : wgs) {
            pl->wgPerXcd = wg;
            if (!timed) break;
            float ms = 1e30f;
            for (int rep = 0; rep < 4 && rc == RSFM_OK; rep++) {
                rc = check_hip(hipEventRecord(e0, pl->stream), "event");
                if (rc == RSFM_OK) rc = enqueue_match(pl, pl->stream);
                if (rc == RSFM_OK) rc = check_hip(hipEventRecord(e1, pl->stream), "event");
                if (rc == RSFM_OK) rc = check_hip(hipEventSynchronize(e1), "event sync");
                float t = 0;
                if (rc == RSFM_OK) rc = check_hip(hipEventElapsedTime(&t, e0, e1), "event time");
                if (rep > 0) ms = std::min(ms, t);  // the first run warms up
            }
            if (rc != RSFM_OK) break;
            pl->tuning.push_back(rsfm_pipeline_tuning_entry{sc, wg, ms});
            if (ms < bestMs) {
                bestMs = ms;
                bestSeg = sc;
                bestWg = wg;
            }
        }
    }
    if (timed) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    RSFM_CHECK(rc);
    pl->tunedMs = timed ? bestMs : 0.f;
    pl->wgPerXcd = bestWg;
    return apply_plan(pl, bestSeg);
}

int capture(rsfm_pipeline* pl)
{
    const auto& c = pl->cfg;
    const uint32_t P = (uint32_t)pl->pairs.size();
    RSFM_CHECK(check_hip(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal), "begin capture"));
    int rc = enqueue_match(pl, pl->stream);
    if (rc == RSFM_OK) rc = aux::launch_filter_table(pl->dFilter, P, c.ratioNum, c.ratioDen, pl->stream);
    if (rc == RSFM_OK && c.matchFilter)
        rc = rmf::launch_batch_table(pl->dRmf, P, pl->rmfCols, pl->rmfRows, pl->rmfParams, pl->stream);
    if (rc == RSFM_OK) rc = undist::launch_pairs_table(pl->dUnd, pl->nUnd, pl->maxUnd, pl->stream);
    if (rc == RSFM_OK && c.nbHyp > 0) rc = ransac::launch_sample_table(pl->dSample, P, c.nbHyp, pl->stream);
    if (c.solver == RSFM_PIPELINE_SOLVER_F8) {
        if (rc == RSFM_OK)
            rc = ransac::launch_score_table(pl->dRansac, P, c.nbHyp, c.threshold, c.mode, c.minSepSq0, c.minSepSq1,
                                            pl->stream, !pl->finish);
        if (rc == RSFM_OK && pl->finish) rc = ransac::launch_finish_table(pl->dFinish, P, pl->finParams, c.mode, pl->stream);
    } else {
        if (rc == RSFM_OK)
            rc = ransac::launch_score7_table(pl->dF7, P, c.nbHyp, c.threshold, c.mode, c.minSepSq0, c.minSepSq1,
                                             pl->stream);
        if (rc == RSFM_OK) rc = degensac::launch_degensac_table(pl->dDg, P, pl->dgParams, c.mode, pl->stream);
        if (rc == RSFM_OK && pl->nPoseF)
            rc = pose::launch_pose_table(pl->dPoseF, pl->nPoseF, 0, c.frontRatio, pl->stream);
        if (rc == RSFM_OK && pl->nPoseH)
            rc = pose::launch_pose_table(pl->dPoseH, pl->nPoseH, 1, c.frontRatio, pl->stream);
    }
    hipGraph_t g = nullptr;
    const int rc2 = check_hip(hipStreamEndCapture(pl->stream, &g), "end capture");
    RSFM_CHECK(rc);
    RSFM_CHECK(rc2);
    pl->graph = g;
    return check_hip(hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0), "graph instantiate");
}

void release(rsfm_pipeline* pl)
{
    if (pl->done) (void)hipEventDestroy(pl->done);
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->arena) (void)hipFree(pl->arena);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

}  // namespace

extern "C" int rsfm_pipeline_create(const rsfm_pipeline_pair* pairs, size_t nbPairs, const rsfm_pipeline_config* cfg,
                                    rsfm_pipeline** out)
{
    RSFM_RETURN_IF(out == nullptr || cfg == nullptr, RSFM_ERR_INVALID, "pipeline: out/cfg is null");
    *out = nullptr;
    RSFM_RETURN_IF(nbPairs == 0 || pairs == nullptr, RSFM_ERR_INVALID, "pipeline: no pairs");
    RSFM_RETURN_IF(cfg->mode != 0 && cfg->mode != 1, RSFM_ERR_INVALID, "pipeline: mode must be 0 or 1");
    RSFM_RETURN_IF(!(cfg->threshold > 0.f), RSFM_ERR_INVALID, "pipeline: threshold must be > 0");
    RSFM_RETURN_IF(cfg->ratioNum == 0 || cfg->ratioDen == 0, RSFM_ERR_INVALID, "pipeline: ratio must be > 0");
    RSFM_RETURN_IF(cfg->segCols % bf::kBN != 0 || cfg->segCols > bf::kMaxSegColsTop2, RSFM_ERR_INVALID,
                   "pipeline: segCols must be 0 or a multiple of 64 up to 4096");
    RSFM_RETURN_IF(cfg->wgPerXcd > 1024 && cfg->wgPerXcd != RSFM_PIPELINE_WG_PER_TILE, RSFM_ERR_INVALID,
                   "pipeline: wgPerXcd must be <= 1024 or RSFM_PIPELINE_WG_PER_TILE");
    RSFM_RETURN_IF(cfg->solver != RSFM_PIPELINE_SOLVER_F8 && cfg->solver != RSFM_PIPELINE_SOLVER_DEGENSAC,
                   RSFM_ERR_INVALID, "pipeline: unknown solver");
    const bool degensac = cfg->solver == RSFM_PIPELINE_SOLVER_DEGENSAC;
    const bool finish = !degensac && (cfg->refit != 0 || cfg->confidence > 0.f);
    ransac::FinishParams finParams{};
    if (finish) RSFM_CHECK(ransac::make_finish_params(cfg->threshold, cfg->mode, cfg->confidence, cfg->maxIter,
                                                      cfg->refit != 0, finParams));
    RSFM_RETURN_IF(cfg->matchFilter > 1, RSFM_ERR_INVALID, "pipeline: matchFilter must be 0 or 1");
    RSFM_RETURN_IF(cfg->matchFilter && (cfg->mfCellCols == 0 || !(cfg->mfThreshold > 0.f)), RSFM_ERR_INVALID,
                   "pipeline: match filter needs mfCellCols >= 1 and mfThreshold > 0");
    for (size_t i = 0; i < nbPairs; i++) {
        const auto& x = pairs[i];
        RSFM_RETURN_IF(x.nA > 0 && (x.descA == nullptr || x.bestA == nullptr || x.secondA == nullptr || x.kpA == nullptr ||
                                    x.pts == nullptr),
                       RSFM_ERR_INVALID, "pipeline pair " + std::to_string(i) + ": query side pointers missing");
        RSFM_RETURN_IF(x.nB > 0 && (x.descB == nullptr || x.bestB == nullptr || x.kpB == nullptr), RSFM_ERR_INVALID,
                       "pipeline pair " + std::to_string(i) + ": train side pointers missing");
        RSFM_RETURN_IF(x.count == nullptr || x.best == nullptr, RSFM_ERR_INVALID,
                       "pipeline pair " + std::to_string(i) + ": count/best missing");
        RSFM_RETURN_IF(cfg->nbHyp > 0 && (x.sampleIdx == nullptr || x.counts == nullptr), RSFM_ERR_INVALID,
                       "pipeline pair " + std::to_string(i) + ": sampleIdx/counts missing");
        RSFM_RETURN_IF(degensac && (x.F == nullptr || x.H == nullptr || x.solverResult == nullptr), RSFM_ERR_INVALID,
                       "pipeline pair " + std::to_string(i) + ": F/H/solverResult missing (DEGENSAC solver)");
        RSFM_RETURN_IF(finish && (x.F == nullptr || x.solverResult == nullptr), RSFM_ERR_INVALID,
                       "pipeline pair " + std::to_string(i) + ": F/solverResult missing (8-point finishing stage)");
        RSFM_RETURN_IF(!degensac && (x.posesF != nullptr || x.posesH != nullptr), RSFM_ERR_INVALID,
                       "pipeline pair " + std::to_string(i) + ": the pose stage needs the DEGENSAC solver");
        RSFM_RETURN_IF((x.posesF != nullptr && x.poseResultF == nullptr) ||
                           (x.posesH != nullptr && x.poseResultH == nullptr),
                       RSFM_ERR_INVALID, "pipeline pair " + std::to_string(i) + ": pose result missing");
        RSFM_RETURN_IF((reinterpret_cast<uintptr_t>(x.descA) | reinterpret_cast<uintptr_t>(x.descB)) % 16 ||
                           (reinterpret_cast<uintptr_t>(x.kpA) | reinterpret_cast<uintptr_t>(x.kpB)) % 8 ||
                           reinterpret_cast<uintptr_t>(x.pts) % 16,
                       RSFM_ERR_INVALID, "pipeline pair " + std::to_string(i) + ": misaligned pointer");
    }
    rsfm_pipeline* pl = new rsfm_pipeline;
    pl->pairs.assign(pairs, pairs + nbPairs);
    pl->cfg = *cfg;
    pl->finish = finish;
    pl->finParams = finParams;
    auto fail_out = [&](int rc) {
        release(pl);
        return rc;
    };
    int rc = check_hip(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking), "pipeline stream");
    if (rc == RSFM_OK) rc = check_hip(hipEventCreateWithFlags(&pl->done, hipEventDisableTiming), "pipeline event");
    if (rc != RSFM_OK) return fail_out(rc);

    // sizes: the finest candidate segment width bounds tiles and partials
    const uint32_t P = (uint32_t)nbPairs;
    const uint32_t minSeg = cfg->segCols ? cfg->segCols : 1024;
    Plan worst = make_plan(pl->pairs, minSeg, nullptr, nullptr);
    pl->tileCap = std::max<size_t>(1, worst.tiles.size());
    pl->partCap = worst.partWords;
    std::vector<pipe::FillRange> fill;
    uint32_t maxFill = 0;
    for (auto& x : pl->pairs) {
        if (x.nA) {
            fill.push_back({reinterpret_cast<uint64_t*>(x.bestA), x.nA});
            fill.push_back({reinterpret_cast<uint64_t*>(x.secondA), x.nA});
        }
        if (x.nB) fill.push_back({reinterpret_cast<uint64_t*>(x.bestB), x.nB});
        maxFill = std::max(maxFill, std::max(x.nA, x.nB));
    }
    pl->nFill = (uint32_t)fill.size();
    pl->maxFill = maxFill;

    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t oTasks = take(P * sizeof(bf::DevTask));
    const size_t oMerge = take(P * sizeof(bf::DevTask));
    const size_t oTiles = take(pl->tileCap * sizeof(bf::TileDesc));
    const size_t oXcd = take((kNumXcd + 1) * sizeof(uint32_t));
    const size_t oFilter = take(P * sizeof(aux::FilterTask));
    const size_t oSample = take(P * sizeof(ransac::SampleTask));
    const size_t oRansac = take(P * sizeof(ransac::DevTask));
    const size_t oFinish = take(P * sizeof(ransac::FinishTask));
    // affine match filter stage: task table and per-pair workspaces
    std::vector<rmf::BatchTask> rt_mf;
    size_t mfBytes = 0;
    if (cfg->matchFilter) {
        rt_mf.resize(P);
        for (uint32_t i = 0; i < P; i++) {
            const auto& x = pl->pairs[i];
            rc = rmf::make_batch_task(rsfm_rmf_task{x.pts, x.matches, x.count, x.nA, nullptr, x.width0, x.height0,
                                                    (uint32_t)(cfg->seed + i)},
                                      cfg->mfCellCols, cfg->mfThreshold, rt_mf[i]);
            if (rc != RSFM_OK) return fail_out(rc);
            mfBytes += rmf::batch_workspace_bytes(x.nA, rt_mf[i].cols * rt_mf[i].rows);
            pl->rmfCols = std::max(pl->rmfCols, rt_mf[i].cols);
            pl->rmfRows = std::max(pl->rmfRows, rt_mf[i].rows);
        }
        pl->rmfParams = rmf::BatchParams{cfg->mfTests, cfg->mfMinVotes, (int)cfg->mfTryOther};
    }
    const size_t oRmf = take(std::max<size_t>(1, rt_mf.size()) * sizeof(rmf::BatchTask));
    const size_t oRmfWs = take(std::max<size_t>(1, mfBytes));
    const size_t oF7 = take(P * sizeof(ransac::DevTask));
    const size_t oDg = take(P * sizeof(degensac::DevTask));
    const size_t oCand = take(degensac ? (size_t)P * cfg->nbCandidates * sizeof(degensac::CandRec) : 1);
    const size_t oUnd = take(P * sizeof(undist::PairTask));
    const size_t oPoseF = take(P * sizeof(pose::DevTask));
    const size_t oPoseH = take(P * sizeof(pose::DevTask));
    const size_t oFill = take(std::max<size_t>(1, fill.size()) * sizeof(pipe::FillRange));
    const size_t oPart = take(std::max<size_t>(1, pl->partCap) * sizeof(uint64_t));
    const size_t oPrep = take(std::max<size_t>(1, worst.prepBytes));
    // 8-point solver: every hypothesis' model, so the best-model kernel reads
    // the winner instead of refitting it on one lane
    const size_t modelFloats = (size_t)cfg->nbHyp * 9;
    const size_t oModels = take(degensac ? 1 : std::max<size_t>(1, (size_t)P * modelFloats * sizeof(float)));
    rc = check_hip(hipMalloc(reinterpret_cast<void**>(&pl->arena), off), "pipeline workspace");
    if (rc != RSFM_OK) return fail_out(rc);
    pl->dTasks = reinterpret_cast<bf::DevTask*>(pl->arena + oTasks);
    pl->dMerge = reinterpret_cast<bf::DevTask*>(pl->arena + oMerge);
    pl->dTiles = reinterpret_cast<bf::TileDesc*>(pl->arena + oTiles);
    pl->dXcdOff = reinterpret_cast<uint32_t*>(pl->arena + oXcd);
    pl->dFilter = reinterpret_cast<aux::FilterTask*>(pl->arena + oFilter);
    pl->dSample = reinterpret_cast<ransac::SampleTask*>(pl->arena + oSample);
    pl->dRansac = reinterpret_cast<ransac::DevTask*>(pl->arena + oRansac);
    pl->dFinish = reinterpret_cast<ransac::FinishTask*>(pl->arena + oFinish);
    pl->dRmf = reinterpret_cast<rmf::BatchTask*>(pl->arena + oRmf);
    {
        uint8_t* ws = pl->arena + oRmfWs;
        for (auto& t : rt_mf) rmf::place_workspace(t, ws);
    }
    pl->dF7 = reinterpret_cast<ransac::DevTask*>(pl->arena + oF7);
    pl->dDg = reinterpret_cast<degensac::DevTask*>(pl->arena + oDg);
    pl->dUnd = reinterpret_cast<undist::PairTask*>(pl->arena + oUnd);
    pl->dPoseF = reinterpret_cast<pose::DevTask*>(pl->arena + oPoseF);
    pl->dPoseH = reinterpret_cast<pose::DevTask*>(pl->arena + oPoseH);
    pl->dFill = reinterpret_cast<pipe::FillRange*>(pl->arena + oFill);
    pl->dPart = reinterpret_cast<uint64_t*>(pl->arena + oPart);
    pl->dPrep = pl->arena + oPrep;
    float* dModels = reinterpret_cast<float*>(pl->arena + oModels);
    for (auto& x : pl->pairs)
        if (x.nA > 0 && x.nB > 0) pl->maxPrepB = std::max(pl->maxPrepB, x.nB);

    std::vector<aux::FilterTask> ft(P);
    std::vector<ransac::SampleTask> st(P);
    std::vector<ransac::DevTask> rt(P), f7;
    std::vector<ransac::FinishTask> fin;
    std::vector<degensac::DevTask> dg;
    std::vector<undist::PairTask> ud;
    std::vector<pose::DevTask> pf, ph;
    if (degensac &&
        (rc = degensac::make_params(cfg->threshold, cfg->thresholdH, cfg->mode, cfg->nbCandidates,
                                    cfg->nbPlaneSamples, cfg->seed, pl->dgParams)) != RSFM_OK)
        return fail_out(rc);
    for (uint32_t i = 0; i < P; i++) {
        const auto& x = pl->pairs[i];
        ft[i] = aux::FilterTask{reinterpret_cast<const uint64_t*>(x.bestA), reinterpret_cast<const uint64_t*>(x.secondA),
                                x.nA, reinterpret_cast<const uint64_t*>(x.bestB), x.nB, x.matches, x.count,
                                reinterpret_cast<const float2*>(x.kpA), reinterpret_cast<const float2*>(x.kpB),
                                reinterpret_cast<float4*>(x.pts)};
        st[i] = ransac::SampleTask{x.sampleIdx, cfg->nbHyp, x.nA, x.count, cfg->seed + i, degensac ? 7u : 8u};
        rt[i] = ransac::DevTask{reinterpret_cast<const float4*>(x.pts), x.count, x.nA, cfg->nbHyp, x.sampleIdx,
                                x.counts, degensac ? nullptr : dModels + i * modelFloats, x.best, x.inlierMask};
        if (finish)
            fin.push_back(ransac::FinishTask{reinterpret_cast<const float4*>(x.pts), x.count, x.nA, x.sampleIdx,
                                             x.counts, dModels + i * modelFloats, cfg->nbHyp, x.F, x.solverResult,
                                             x.inlierMask, x.best});
        if (x.distortion) {
            rsfm_real_camera c0{x.cam0, {}}, c1{x.cam1, {}};
            for (int k = 0; k < 5; k++) {
                c0.distortion[k] = x.distortion[k];
                c1.distortion[k] = x.distortion[5 + k];
            }
            undist::PairTask u{};
            bool a0 = false, a1 = false;
            if ((rc = undist::make_params(c0, 2, u.d0, a0)) != RSFM_OK || (rc = undist::make_params(c1, 2, u.d1, a1)) != RSFM_OK)
                return fail_out(rc);
            if (a0 || a1) {
                u.pts = reinterpret_cast<float4*>(x.pts);
                u.nDev = x.count;
                u.nHost = x.nA;
                u.on0 = a0;
                u.on1 = a1;
                ud.push_back(u);
                pl->maxUnd = std::max(pl->maxUnd, x.nA);
            }
        }
        if (!degensac) continue;
        f7.push_back(ransac::DevTask{reinterpret_cast<const float4*>(x.pts), x.count, x.nA, cfg->nbHyp, x.sampleIdx,
                                     x.counts, nullptr, x.best, nullptr});
        dg.emplace_back();
        if ((rc = degensac::make_task(rsfm_degensac_task{x.pts, x.nA, x.count, x.sampleIdx, x.counts, cfg->nbHyp, x.F,
                                                         x.H, x.solverResult, x.inlierMask, x.inlierMaskH},
                                      dg.back())) != RSFM_OK)
            return fail_out(rc);
        dg.back().cand = reinterpret_cast<degensac::CandRec*>(pl->arena + oCand) + (size_t)(dg.size() - 1) * cfg->nbCandidates;
        if (x.posesF) {
            pf.emplace_back();
            if ((rc = pose::make_task(rsfm_pose_task{x.F, x.cam0, x.cam1, x.pts, x.nA, x.count, x.inlierMask,
                                                     x.posesF, x.poseResultF},
                                      pf.back())) != RSFM_OK)
                return fail_out(rc);
        }
        if (x.posesH) {
            ph.emplace_back();
            if ((rc = pose::make_task(rsfm_pose_h_task{x.H, x.cam0, x.cam1, x.pts, x.nA, x.count, x.inlierMaskH,
                                                       x.posesH, x.normalsH, x.poseResultH},
                                      ph.back())) != RSFM_OK)
                return fail_out(rc);
        }
    }
    if ((rc = upload(pl->dFilter, ft.data(), P * sizeof(aux::FilterTask))) != RSFM_OK ||
        (rc = upload(pl->dSample, st.data(), P * sizeof(ransac::SampleTask))) != RSFM_OK ||
        (rc = upload(pl->dRansac, rt.data(), P * sizeof(ransac::DevTask))) != RSFM_OK ||
        (rc = upload(pl->dFinish, fin.data(), fin.size() * sizeof(ransac::FinishTask))) != RSFM_OK ||
        (rc = upload(pl->dRmf, rt_mf.data(), rt_mf.size() * sizeof(rmf::BatchTask))) != RSFM_OK ||
        (rc = upload(pl->dF7, f7.data(), f7.size() * sizeof(ransac::DevTask))) != RSFM_OK ||
        (rc = upload(pl->dDg, dg.data(), dg.size() * sizeof(degensac::DevTask))) != RSFM_OK ||
        (rc = upload(pl->dUnd, ud.data(), ud.size() * sizeof(undist::PairTask))) != RSFM_OK ||
        (rc = upload(pl->dPoseF, pf.data(), pf.size() * sizeof(pose::DevTask))) != RSFM_OK ||
        (rc = upload(pl->dPoseH, ph.data(), ph.size() * sizeof(pose::DevTask))) != RSFM_OK ||
        (rc = upload(pl->dFill, fill.data(), fill.size() * sizeof(pipe::FillRange))) != RSFM_OK)
        return fail_out(rc);
    pl->nUnd = (uint32_t)ud.size();
    pl->nPoseF = (uint32_t)pf.size();
    pl->nPoseH = (uint32_t)ph.size();
    if (degensac && (rc = pose::check_ratio(cfg->frontRatio)) != RSFM_OK) return fail_out(rc);
    if ((rc = tune(pl)) != RSFM_OK) return fail_out(rc);
    if ((rc = capture(pl)) != RSFM_OK) return fail_out(rc);
    *out = pl;
    return RSFM_OK;
}

extern "C" int rsfm_pipeline_run(rsfm_pipeline* pl, hipStream_t stream)
{
    RSFM_RETURN_IF(pl == nullptr || pl->exec == nullptr, RSFM_ERR_INVALID, "pipeline: not created");
    RSFM_CHECK(check_hip(hipGraphLaunch(pl->exec, stream), "pipeline graph launch"));
    pl->ran = true;
    return check_hip(hipEventRecord(pl->done, stream), "pipeline run event");
}

extern "C" int rsfm_pipeline_info(const rsfm_pipeline* pl, rsfm_pipeline_info_t* info, rsfm_pipeline_tuning_entry* entries,
                                  size_t maxEntries, size_t* nbEntries)
{
    RSFM_RETURN_IF(pl == nullptr || info == nullptr, RSFM_ERR_INVALID, "pipeline info: null argument");
    info->segCols = pl->segCols;
    info->wgPerXcd = pl->wgPerXcd;
    info->nbTiles = pl->nTiles;
    info->tunedMatchMs = pl->tunedMs;
    if (nbEntries) *nbEntries = pl->tuning.size();
    if (entries)
        for (size_t i = 0; i < std::min(maxEntries, pl->tuning.size()); i++) entries[i] = pl->tuning[i];
    return RSFM_OK;
}

extern "C" int rsfm_pipeline_destroy(rsfm_pipeline* pl)
{
    if (pl == nullptr) return RSFM_OK;
    // the graph's last run may still read the arena on the caller's stream
    int rc = pl->ran ? check_hip(hipEventSynchronize(pl->done), "pipeline run sync") : RSFM_OK;
    const int rc2 = check_hip(hipStreamSynchronize(pl->stream), "pipeline sync");
    if (rc == RSFM_OK) rc = rc2;
    release(pl);
    return rc;
}
