// rsfm_amd.hpp - C++ adapters with the reference's interfaces over the C ABI.
//
// A RapidSFM build switches from its CUDA kernels to the MI355X library by
// including this header in place of bruteForceMatch.h, casHash/cascadeHash.h,
// RansacMatchFilter.cuh/.h and bruteForceMatcher.h: the function names,
// argument order and meaning are the reference's, streams are hipStream_t, and
// failures throw rsfm::Exception (the reference throws cudapp::Exception from
// cudaCheck). Header-only; link against librsfm_amd.so.
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <future>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "rsfm_amd.h"

namespace rsfm {

struct Exception : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void rsfmCheck(int rc, const char* what)
{
    if (rc != RSFM_OK) throw Exception(std::string(what) + ": " + rsfm_last_error());
}
inline void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Exception(std::string(what) + ": " + hipGetErrorString(e));
}

// ---- bruteForceMatch.h:76-136 ----------------------------------------------
using BestMatch = rsfm_best_match;
using ValidMatch = rsfm_valid_match;
using SiftBruteForceMatchTask = rsfm_bf_task;
// same 56-B layout; a distinct type so the 4-bit overload resolves as in the reference
struct Sift4bBruteForceMatchTask : rsfm_bf_task {};
static_assert(sizeof(SiftBruteForceMatchTask) == 56 && sizeof(Sift4bBruteForceMatchTask) == 56, "task layout");
static_assert(sizeof(BestMatch) == 8 && sizeof(ValidMatch) == 12, "result layout");

inline void launchPreCompSiftSqrNorm(uint32_t* sqrNorm, const void* src, uint32_t nbDesc, hipStream_t stream)
{
    rsfmCheck(rsfm_sqrnorm_u8(sqrNorm, src, nbDesc, stream), "launchPreCompSiftSqrNorm");
}
inline void launchBruteForceMatchIMMA(const SiftBruteForceMatchTask* tasks, size_t nbTasks, hipStream_t stream)
{
    rsfmCheck(rsfm_bf_match_u8(tasks, nbTasks, stream), "launchBruteForceMatchIMMA");
}
inline void launchBruteForceMatchIDP(const SiftBruteForceMatchTask* tasks, size_t nbTasks, hipStream_t stream)
{
    rsfmCheck(rsfm_bf_match_idp_u8(tasks, nbTasks, stream), "launchBruteForceMatchIDP");
}
inline void launchSiftDesc8bTo4b(void* dst, const void* src, uint32_t nbDesc, hipStream_t stream)
{
    rsfmCheck(rsfm_desc8b_to_4b(dst, src, nbDesc, stream), "launchSiftDesc8bTo4b");
}
inline void launchPreCompSift4bSqrNorm(uint32_t* sqrNorm, const void* src, uint32_t nbDesc, hipStream_t stream)
{
    rsfmCheck(rsfm_sqrnorm_u4(sqrNorm, src, nbDesc, stream), "launchPreCompSift4bSqrNorm");
}
inline void launchBruteForceMatchIMMA(const Sift4bBruteForceMatchTask* tasks, size_t nbTasks, hipStream_t stream)
{
    std::vector<rsfm_bf_task> t(tasks, tasks + nbTasks);
    rsfmCheck(rsfm_bf_match_u4(t.data(), nbTasks, stream), "launchBruteForceMatchIMMA(4b)");
}

// bruteForceMatch.cpp:23-66
inline std::vector<ValidMatch> crossCheckMatches(const BestMatch* matches, uint32_t nbQueryDesc, const BestMatch* matchesBwd,
                                                 uint32_t nbTrainDesc)
{
    std::vector<ValidMatch> out(std::min(nbQueryDesc, nbTrainDesc));
    out.resize(rsfm_cross_check_host(matches, nbQueryDesc, matchesBwd, nbTrainDesc, out.data()));
    return out;
}
inline std::vector<ValidMatch> removeMatchConflicts(const BestMatch* matches, uint32_t nbQueryDesc, uint32_t nbTrainDesc)
{
    std::vector<ValidMatch> out(std::min(nbQueryDesc, nbTrainDesc));
    out.resize(rsfm_remove_conflicts_host(matches, nbQueryDesc, nbTrainDesc, out.data()));
    return out;
}

// ---- casHash/cascadeHash.h:23-107 ------------------------------------------
namespace cashash {
static constexpr uint32_t nbScheme = 8;
static constexpr uint32_t nbBuckets = 256;
using ImageDesc = rsfm_cashash_image;
using MatchResult = rsfm_cashash_result;
static_assert(sizeof(ImageDesc) == 32 && sizeof(MatchResult) == 16, "cashash layouts");

// acc / sqrAcc: device uint64_t[128]; weights: device int8[64][128]
inline void accumuateDesc(uint64_t* acc, uint64_t* sqrAcc, bool initAcc, const void* desc, uint32_t nbDesc,
                          hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_accumulate_desc(acc, sqrAcc, initAcc, desc, nbDesc, stream), "cashash::accumuateDesc");
}
inline void computeBias(int32_t* bias, const int8_t* weights, const int32_t* biasBase, const uint64_t* acc, uint64_t nbDesc,
                        hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_compute_bias(bias, weights, biasBase, acc, nbDesc, 64, stream), "cashash::computeBias");
}
inline void computeHashForBucketAndRemap(const int8_t* weights, const int32_t* bias, const void* desc, uint32_t nbDesc,
                                         uint8_t* bucketIndices, uint32_t* remap, uint32_t* histogram, hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_compute_hash_bucket_and_remap(weights, bias, desc, nbDesc, bucketIndices, remap, histogram, stream),
              "cashash::computeHashForBucketAndRemap");
}
inline void computeHashForBucketOnly(const int8_t* weights, const int32_t* bias, const void* desc, uint32_t nbDesc,
                                     uint8_t* bucketIndices, uint32_t* histogram, hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_compute_hash_bucket_only(weights, bias, desc, nbDesc, bucketIndices, histogram, stream),
              "cashash::computeHashForBucketOnly");
}
inline void buildHistogram(uint32_t* histogram, const uint8_t* bucketIndices, uint32_t nbDesc, hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_build_histogram(histogram, bucketIndices, nbDesc, stream), "cashash::buildHistogram");
}
inline void computeBucketBound(uint32_t* lower, uint32_t* upper, uint32_t* maxBucketSize, const uint32_t* histogram,
                               hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_compute_bucket_bound(lower, upper, maxBucketSize, histogram, stream), "cashash::computeBucketBound");
}
inline void buildBuckets(const uint8_t* bucketIndices, uint32_t nbDesc, uint32_t* idxAllocator, uint32_t* buckets,
                         hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_build_buckets(bucketIndices, nbDesc, idxAllocator, buckets, stream), "cashash::buildBuckets");
}
// Overloads with the reference's own parameter types (casHash/cascadeHash.h:32-67,
// called as at bruteForceMatcher.cpp:97-103): the reference passes
// cudapp::KArray objects in device memory by reference or pointer. Any type of
// the same size and trivial layout binds here (cudapp's KArray is a plain
// array wrapper): KArray<uint64_t,128> (acc), KArray<int32_t,64> (bias, bias
// base), KArray<int32_t,64,32> (weights = int8 [64][128]), KArray<int32_t,32>
// (one weight row), KArray<uint32_t,32> (descriptor), KArray<uint8_t,8>
// (bucket codes), KArray<uint32_t,8,256> (histogram, bounds, allocator).
namespace detail {
template <class T, size_t kBytes>
using block_t = std::enable_if_t<std::is_trivially_copyable<T>::value && sizeof(T) == kBytes, int>;
inline std::nullptr_t dev(std::nullptr_t) { return nullptr; }
template <class T>
inline T* dev(T* p) { return p; }
}  // namespace detail

template <class Acc, class Desc, class SqrAcc, detail::block_t<Acc, 1024> = 0, detail::block_t<Desc, 128> = 0>
inline void accumuateDesc(Acc& acc, SqrAcc* sqrAcc, bool initAcc, const Desc* desc, uint32_t nbDesc, hipStream_t stream)
{
    static_assert(sizeof(SqrAcc) == 1024, "sqrAcc: KArray<uint64_t, 128>");
    accumuateDesc(reinterpret_cast<uint64_t*>(&acc), reinterpret_cast<uint64_t*>(sqrAcc), initAcc, desc, nbDesc, stream);
}
template <class Acc, class Desc, detail::block_t<Acc, 1024> = 0, detail::block_t<Desc, 128> = 0>
inline void accumuateDesc(Acc& acc, std::nullptr_t, bool initAcc, const Desc* desc, uint32_t nbDesc, hipStream_t stream)
{
    accumuateDesc(reinterpret_cast<uint64_t*>(&acc), nullptr, initAcc, desc, nbDesc, stream);
}
template <class Bias, class W, class Base, class Acc, detail::block_t<Bias, 256> = 0, detail::block_t<W, 8192> = 0,
          detail::block_t<Base, 256> = 0, detail::block_t<Acc, 1024> = 0>
inline void computeBias(Bias& bias, const W& weights, const Base& biasBase, const Acc& acc, uint64_t nbDesc,
                        hipStream_t stream)
{
    computeBias(reinterpret_cast<int32_t*>(&bias), reinterpret_cast<const int8_t*>(&weights),
                reinterpret_cast<const int32_t*>(&biasBase), reinterpret_cast<const uint64_t*>(&acc), nbDesc, stream);
}
template <class WRow, class Desc, class Codes, class Hist, detail::block_t<WRow, 128> = 0, detail::block_t<Desc, 128> = 0,
          detail::block_t<Codes, 8> = 0>
inline void computeHashForBucketOnly(const WRow* weights, const int32_t* bias, const Desc* desc, uint32_t nbDesc,
                                     Codes* bucketIndices, Hist* histogram, hipStream_t stream)
{
    static_assert(std::is_same<Hist, std::nullptr_t>::value || sizeof(Hist) == 8192, "histogram: KArray<uint32_t, 8, 256>");
    computeHashForBucketOnly(reinterpret_cast<const int8_t*>(weights), bias, desc, nbDesc,
                             reinterpret_cast<uint8_t*>(bucketIndices), reinterpret_cast<uint32_t*>(histogram), stream);
}
template <class WRow, class Desc, class Codes, detail::block_t<WRow, 128> = 0, detail::block_t<Desc, 128> = 0,
          detail::block_t<Codes, 8> = 0>
inline void computeHashForBucketOnly(const WRow* weights, const int32_t* bias, const Desc* desc, uint32_t nbDesc,
                                     Codes* bucketIndices, std::nullptr_t, hipStream_t stream)
{
    computeHashForBucketOnly(reinterpret_cast<const int8_t*>(weights), bias, desc, nbDesc,
                             reinterpret_cast<uint8_t*>(bucketIndices), nullptr, stream);
}
template <class WRow, class Desc, class Codes, class Remap, detail::block_t<WRow, 128> = 0,
          detail::block_t<Desc, 128> = 0, detail::block_t<Codes, 8> = 0, detail::block_t<Remap, 16> = 0>
inline void computeHashForBucketAndRemap(const WRow* weights, const int32_t* bias, const Desc* desc, uint32_t nbDesc,
                                         Codes* bucketIndices, Remap* remap, std::nullptr_t, hipStream_t stream)
{
    computeHashForBucketAndRemap(reinterpret_cast<const int8_t*>(weights), bias, desc, nbDesc,
                                 reinterpret_cast<uint8_t*>(bucketIndices), reinterpret_cast<uint32_t*>(remap), nullptr,
                                 stream);
}
template <class Hist, class Codes, detail::block_t<Hist, 8192> = 0, detail::block_t<Codes, 8> = 0>
inline void buildHistogram(Hist* histogram, const Codes* bucketIndices, uint32_t nbDesc, hipStream_t stream)
{
    buildHistogram(reinterpret_cast<uint32_t*>(histogram), reinterpret_cast<const uint8_t*>(bucketIndices), nbDesc, stream);
}
template <class Hist, detail::block_t<Hist, 8192> = 0>
inline void computeBucketBound(Hist& lower, Hist& upper, uint32_t* maxBucketSize, const Hist& histogram,
                               hipStream_t stream)
{
    computeBucketBound(reinterpret_cast<uint32_t*>(&lower), reinterpret_cast<uint32_t*>(&upper), maxBucketSize,
                       reinterpret_cast<const uint32_t*>(&histogram), stream);
}
template <class Codes, class Hist, detail::block_t<Codes, 8> = 0, detail::block_t<Hist, 8192> = 0>
inline void buildBuckets(const Codes* bucketIndices, uint32_t nbDesc, Hist& idxAllocator, uint32_t* buckets,
                         hipStream_t stream)
{
    buildBuckets(reinterpret_cast<const uint8_t*>(bucketIndices), nbDesc, reinterpret_cast<uint32_t*>(&idxAllocator),
                 buckets, stream);
}
inline void findBestMatch(const ImageDesc* queries, const uint32_t* querySize, const ImageDesc& train, uint32_t trainSize,
                          const MatchResult* results, uint32_t nbTasks, const uint32_t* maxQueryBucketSize,
                          uint32_t maxTrainBucketSize, hipStream_t stream)
{
    rsfmCheck(rsfm_cashash_find_best_match(queries, querySize, &train, trainSize, results, nbTasks, maxQueryBucketSize,
                                           maxTrainBucketSize, nbScheme, stream),
              "cashash::findBestMatch");
}
}  // namespace cashash

// ---- RansacMatchFilter.cuh:21-34 -------------------------------------------
namespace rmf {
using PtPair = rsfm_pt_pair;
// kmat<float,2,3> per cell is 6 floats
inline void cudaRansacMatchFilter(uint8_t* votes, float* affineScratch, const PtPair* matches, uint32_t nbMatches,
                                  const uint32_t* bounds, uint32_t cols, uint32_t rows, uint32_t nbRansacTests,
                                  float threshold, bool tryOtherSolutions, hipStream_t stream, uint32_t seed = 0)
{
    rsfmCheck(rsfm_rmf_filter(votes, affineScratch, matches, nbMatches, bounds, cols, rows, nbRansacTests, threshold,
                              tryOtherSolutions, seed, stream),
              "rmf::cudaRansacMatchFilter");
}
}  // namespace rmf

// ---- device buffer used by the two operator classes -------------------------
template <typename T>
struct DeviceBuffer {
    T* ptr = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { reserve(n); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        std::swap(ptr, o.ptr);
        std::swap(count, o.count);
        return *this;
    }
    ~DeviceBuffer()
    {
        if (ptr) (void)hipFree(ptr);
    }
    // grows to at least n elements; true when it (re)allocated
    bool reserve(size_t n)
    {
        if (n <= count && ptr) return false;
        if (ptr) hipCheck(hipFree(ptr), "hipFree");
        ptr = nullptr;
        hipCheck(hipMalloc(&ptr, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
        count = std::max<size_t>(n, 1);
        return true;
    }
};

// page-locked host memory (asynchronous device-to-host copies)
template <typename T>
struct PinnedBuffer {
    T* ptr = nullptr;
    size_t count = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    PinnedBuffer(PinnedBuffer&& o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
    ~PinnedBuffer()
    {
        if (ptr) (void)hipHostFree(ptr);
    }
    bool reserve(size_t n)
    {
        if (n <= count && ptr) return false;
        if (ptr) hipCheck(hipHostFree(ptr), "hipHostFree");
        ptr = nullptr;
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault),
                 "hipHostMalloc");
        count = std::max<size_t>(n, 1);
        return true;
    }
};

// ---- RansacMatchFilter.h:28-44 -----------------------------------------------
using Vec2f = std::array<float, 2>;

class RansacMatchFilter {
public:
    RansacMatchFilter() = default;
    // threshold is relative to the cell width (RansacMatchFilter.h:34); seed
    // replaces the reference's clock() seed
    std::vector<bool> getInlierMask(std::array<int, 2> imgSize0, const std::vector<std::pair<Vec2f, Vec2f>>& matches,
                                    uint32_t minVotes = 2, uint32_t nbRansacTests = 256, float threshold = 0.15f,
                                    uint32_t cellCols = 16, bool tryOtherAffines = false, hipStream_t stream = nullptr,
                                    uint32_t seed = 0)
    {
        std::lock_guard<std::mutex> lk(mMutex);
        const uint32_t n = static_cast<uint32_t>(matches.size());
        std::vector<rsfm_pt_pair> in(n);
        for (uint32_t k = 0; k < n; k++)
            in[k] = rsfm_pt_pair{matches[k].first[0], matches[k].first[1], matches[k].second[0], matches[k].second[1]};
        const uint32_t cw0 = (uint32_t(std::max(imgSize0[0], imgSize0[1])) + cellCols - 1) / cellCols;
        const uint32_t nCells = ((uint32_t(imgSize0[0]) + cw0 - 1) / cw0) * ((uint32_t(imgSize0[1]) + cw0 - 1) / cw0);
        std::vector<rsfm_pt_pair> reordered(n);
        std::vector<uint32_t> bounds(nCells + 1), order(n);
        uint32_t cols = 0, rows = 0, cw = 0;
        rsfmCheck(rsfm_rmf_bin_host(imgSize0[0], imgSize0[1], cellCols, in.data(), n, &cols, &rows, &cw, reordered.data(),
                                    bounds.data(), order.data()),
                  "RansacMatchFilter binning");
        const size_t voteBytes = (n + 3) / 4 * 4;
        mVotes.reserve(voteBytes);
        mAffine.reserve(size_t(nCells) * 6);
        mMatches.reserve(n);
        mBounds.reserve(nCells + 1);
        hipCheck(hipMemcpyAsync(mMatches.ptr, reordered.data(), n * sizeof(rsfm_pt_pair), hipMemcpyHostToDevice, stream),
                 "H2D matches");
        hipCheck(hipMemcpyAsync(mBounds.ptr, bounds.data(), bounds.size() * 4, hipMemcpyHostToDevice, stream), "H2D bounds");
        rmf::cudaRansacMatchFilter(mVotes.ptr, mAffine.ptr, mMatches.ptr, n, mBounds.ptr, cols, rows, nbRansacTests,
                                   float(cw) * threshold, tryOtherAffines, stream, seed);
        std::vector<uint8_t> votes(n);
        hipCheck(hipMemcpyAsync(votes.data(), mVotes.ptr, n, hipMemcpyDeviceToHost, stream), "D2H votes");
        hipCheck(hipStreamSynchronize(stream), "sync");
        std::vector<bool> mask(n);
        for (uint32_t i = 0; i < n; i++) mask[order[i]] = votes[i] >= minVotes;
        return mask;
    }

private:
    std::mutex mMutex;
    DeviceBuffer<uint8_t> mVotes;
    DeviceBuffer<float> mAffine;
    DeviceBuffer<rsfm_pt_pair> mMatches;
    DeviceBuffer<uint32_t> mBounds;
};

// ---- bruteForceMatcher.h:28-102 ----------------------------------------------
using ImageHandle = uint32_t;

// Per-image precompute cache and batched matching of queries against one
// train image (bruteForceMatcher.cpp:163-443), as a GPU-resident service:
//  - result storage per query slot (device, plus page-locked host memory for
//    the copies back) grows by capacity and is reused across calls, as the
//    reference's reserveStorage does (bruteForceMatcher.cpp:209-237); the
//    private stream is synchronised only before storage grows;
//  - results come back through page-locked memory and a host function enqueued
//    on the private stream fulfils the promise (bruteForceMatcher.cpp:360-396),
//    so no call blocks the host;
//  - registration keeps its scratch (bias, codes, histogram, bounds) across
//    calls and does not synchronise; the cascade-hash method keeps the running
//    descriptor mean in registration order (bruteForceMatcher.cpp:97-100).
class BruteForceMatcher {
public:
    using Descriptor = std::array<uint32_t, 32>;
    struct Config {
        bool useCascadeHash = false;
        bool useSift4b = false;
        // cascade hash projection (int8 [64][128]) and bias base, host memory;
        // the reference draws them from default_random_engine{11}
        std::vector<int8_t> casHashWeights;
        std::vector<int32_t> casHashBiasBase;
    };
    struct Img {
        const Descriptor* desc;  // device
        uint32_t nbDesc;
        ImageHandle hImage;
    };
    using Result = std::vector<std::pair<std::vector<BestMatch>, std::vector<BestMatch>>>;

    BruteForceMatcher() : BruteForceMatcher(Config()) {}
    explicit BruteForceMatcher(Config cfg) : mCfg(std::move(cfg))
    {
        hipCheck(hipStreamCreate(&mStream), "hipStreamCreate");
        if (mCfg.useCascadeHash) {
            if (mCfg.casHashWeights.size() != 64 * 128 || mCfg.casHashBiasBase.size() != 64)
                throw Exception("BruteForceMatcher: cascade hash needs 64x128 weights and 64 bias values");
            grow(mWeights, 64 * 128);
            grow(mBiasBase, 64);
            grow(mAcc, 128);
            grow(mBias, 64);
            hipCheck(hipMemcpy(mWeights.ptr, mCfg.casHashWeights.data(), 64 * 128, hipMemcpyHostToDevice), "H2D weights");
            hipCheck(hipMemcpy(mBiasBase.ptr, mCfg.casHashBiasBase.data(), 64 * 4, hipMemcpyHostToDevice), "H2D bias base");
        }
    }
    BruteForceMatcher(const BruteForceMatcher&) = delete;
    BruteForceMatcher& operator=(const BruteForceMatcher&) = delete;
    ~BruteForceMatcher()
    {
        if (mStream) {
            (void)hipStreamSynchronize(mStream);  // pending copies and host functions use the storage
            (void)hipStreamDestroy(mStream);
        }
    }

    void registerImage(ImageHandle hImage, const Descriptor* desc, uint32_t nbDesc, hipStream_t userStream)
    {
        std::lock_guard<std::mutex> lk(mLock);
        join(userStream, mStream);
        PreComp pc;
        pc.nbDesc = nbDesc;
        if (mCfg.useSift4b) {
            grow(pc.desc4, size_t(nbDesc) * 64);
            launchSiftDesc8bTo4b(pc.desc4.ptr, desc, nbDesc, mStream);
        }
        if (mCfg.useCascadeHash) {
            const bool first = mAccNbDesc == 0;
            cashash::accumuateDesc(mAcc.ptr, nullptr, first, desc, nbDesc, mStream);
            mAccNbDesc += nbDesc;
            cashash::computeBias(mBias.ptr, mWeights.ptr, mBiasBase.ptr, mAcc.ptr, mAccNbDesc, mStream);
            if (mCodes.count < size_t(nbDesc) * 8 || mHist.count < 2048 || mLower.count < 2048) {
                hipCheck(hipStreamSynchronize(mStream), "registerImage: grow scratch");
                grow(mCodes, size_t(nbDesc) * 8);
                grow(mHist, 2048);
                grow(mLower, 2048);
            }
            grow(pc.ends, 2048);
            grow(pc.buckets, size_t(nbDesc) * 8);
            cashash::computeHashForBucketOnly(mWeights.ptr, mBias.ptr, desc, nbDesc, mCodes.ptr, nullptr, mStream);
            cashash::buildHistogram(mHist.ptr, mCodes.ptr, nbDesc, mStream);
            cashash::computeBucketBound(mLower.ptr, pc.ends.ptr, nullptr, mHist.ptr, mStream);
            cashash::buildBuckets(mCodes.ptr, nbDesc, mLower.ptr, pc.buckets.ptr, mStream);
        }
        mPreComp[hImage] = std::move(pc);
        join(mStream, userStream);
    }

    std::future<Result> match(const std::vector<Img>& queries, const Img& train, bool bidirectional, hipStream_t userStream)
    {
        std::lock_guard<std::mutex> lk(mLock);
        (void)bidirectional;  // bidirectional is forced on, as in bruteForceMatcher.cpp:207
        auto job = new MatchJob;
        auto fut = job->promise.get_future();
        if (queries.empty()) {
            job->promise.set_value(Result{});
            delete job;
            return fut;
        }
        join(userStream, mStream);
        matchImpl(queries, train);
        for (size_t i = 0; i < queries.size(); i++) {
            job->slots.push_back({mFwd[i].host.ptr, queries[i].nbDesc, mBwd[i].host.ptr, train.nbDesc});
            if (queries[i].nbDesc)
                hipCheck(hipMemcpyAsync(mFwd[i].host.ptr, mFwd[i].dev.ptr, queries[i].nbDesc * sizeof(BestMatch),
                                        hipMemcpyDeviceToHost, mStream), "D2H fwd");
            if (train.nbDesc)
                hipCheck(hipMemcpyAsync(mBwd[i].host.ptr, mBwd[i].dev.ptr, train.nbDesc * sizeof(BestMatch),
                                        hipMemcpyDeviceToHost, mStream), "D2H bwd");
        }
        hipCheck(hipLaunchHostFunc(mStream, &MatchJob::run, job), "hipLaunchHostFunc");
        join(mStream, userStream);
        return fut;
    }

    // With crossCheck the filter runs on the device (rsfm_filter_matches_batch):
    // the kept matches and their counts come back through page-locked memory,
    // where the reference copies both full best-match arrays and filters on the
    // CPU (bruteForceMatcher.cpp:354-369). Without crossCheck,
    // removeMatchConflicts runs on the host as in the reference.
    std::future<std::vector<std::vector<ValidMatch>>> matchAndFilter(const std::vector<Img>& queries, const Img& train,
                                                                    bool crossCheck, hipStream_t userStream)
    {
        if (!crossCheck) {
            auto f = std::make_shared<std::future<Result>>(match(queries, train, true, userStream));
            const uint32_t nT = train.nbDesc;
            return std::async(std::launch::deferred, [f, nT]() {
                Result r = f->get();
                std::vector<std::vector<ValidMatch>> out(r.size());
                for (size_t i = 0; i < r.size(); i++)
                    out[i] = removeMatchConflicts(r[i].first.data(), uint32_t(r[i].first.size()), nT);
                return out;
            });
        }
        std::lock_guard<std::mutex> lk(mLock);
        auto job = new FilterJob;
        auto fut = job->promise.get_future();
        const size_t nq = queries.size();
        if (nq == 0) {
            job->promise.set_value({});
            delete job;
            return fut;
        }
        join(userStream, mStream);
        matchImpl(queries, train);
        bool regrow = mKept.size() < nq || mCounts.count < nq || mHostCounts.count < nq;
        for (size_t i = 0; i < std::min(nq, mKept.size()); i++)
            regrow |= mKept[i].dev.count < queries[i].nbDesc || mKept[i].host.count < queries[i].nbDesc;
        if (regrow) {
            hipCheck(hipStreamSynchronize(mStream), "matchAndFilter: grow storage");
            if (mKept.size() < nq) mKept.resize(nq);
            grow(mCounts, nq);
            grow(mHostCounts, nq);
            for (size_t i = 0; i < nq; i++) {
                grow(mKept[i].dev, queries[i].nbDesc);
                grow(mKept[i].host, queries[i].nbDesc);
            }
        }
        std::vector<rsfm_filter_task> tasks(nq);
        for (size_t i = 0; i < nq; i++)
            tasks[i] = rsfm_filter_task{mFwd[i].dev.ptr, nullptr, queries[i].nbDesc, mBwd[i].dev.ptr, train.nbDesc,
                                        mKept[i].dev.ptr, mCounts.ptr + i, nullptr, nullptr, nullptr};
        rsfmCheck(rsfm_filter_matches_batch(tasks.data(), nq, 4, 5, mStream), "rsfm_filter_matches_batch");
        hipCheck(hipMemcpyAsync(mHostCounts.ptr, mCounts.ptr, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, mStream),
                 "D2H counts");
        for (size_t i = 0; i < nq; i++) {
            if (queries[i].nbDesc)
                hipCheck(hipMemcpyAsync(mKept[i].host.ptr, mKept[i].dev.ptr, queries[i].nbDesc * sizeof(ValidMatch),
                                        hipMemcpyDeviceToHost, mStream),
                         "D2H kept matches");
            job->kept.push_back(mKept[i].host.ptr);
        }
        job->counts = mHostCounts.ptr;
        hipCheck(hipLaunchHostFunc(mStream, &FilterJob::run, job), "hipLaunchHostFunc");
        join(mStream, userStream);
        return fut;
    }

    // device and page-locked allocations made so far (storage, scratch and
    // per-image precompute data); constant across calls once storage is large
    // enough
    size_t allocations() const { return mAllocations; }

private:
    struct PreComp {
        uint32_t nbDesc = 0;
        DeviceBuffer<uint8_t> desc4;
        DeviceBuffer<uint32_t> ends;
        DeviceBuffer<uint32_t> buckets;
    };
    struct MatchStorage {
        DeviceBuffer<BestMatch> dev;
        PinnedBuffer<BestMatch> host;
    };
    struct KeptStorage {
        DeviceBuffer<ValidMatch> dev;
        PinnedBuffer<ValidMatch> host;
    };
    // results of one match() call, filled from page-locked memory by a host
    // function that runs after the copies on the private stream
    struct MatchJob {
        struct Slot {
            const BestMatch* fwd;
            uint32_t nFwd;
            const BestMatch* bwd;
            uint32_t nBwd;
        };
        std::vector<Slot> slots;
        std::promise<Result> promise;
        static void run(void* p)
        {
            std::unique_ptr<MatchJob> job(static_cast<MatchJob*>(p));
            Result r(job->slots.size());
            for (size_t i = 0; i < job->slots.size(); i++) {
                const Slot& s = job->slots[i];
                r[i].first.assign(s.fwd, s.fwd + s.nFwd);
                r[i].second.assign(s.bwd, s.bwd + s.nBwd);
            }
            job->promise.set_value(std::move(r));
        }
    };
    struct FilterJob {
        std::vector<const ValidMatch*> kept;
        const uint32_t* counts = nullptr;
        std::promise<std::vector<std::vector<ValidMatch>>> promise;
        static void run(void* p)
        {
            std::unique_ptr<FilterJob> job(static_cast<FilterJob*>(p));
            std::vector<std::vector<ValidMatch>> out(job->kept.size());
            for (size_t i = 0; i < out.size(); i++) out[i].assign(job->kept[i], job->kept[i] + job->counts[i]);
            job->promise.set_value(std::move(out));
        }
    };

    template <class Buf>
    void grow(Buf& b, size_t n)
    {
        if (b.reserve(n)) mAllocations++;
    }

    static void join(hipStream_t from, hipStream_t to)
    {
        if (from == to) return;
        hipEvent_t e;
        hipCheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        hipCheck(hipEventRecord(e, from), "record");
        hipCheck(hipStreamWaitEvent(to, e, 0), "wait");
        hipCheck(hipEventDestroy(e), "destroy");
    }

    // query slot i's storage holds at least nbDesc (fwd) and the train size (bwd);
    // the private stream is drained before anything is reallocated
    void reserveStorage(const std::vector<Img>& queries, const Img& train)
    {
        bool regrow = mFwd.size() < queries.size();
        for (size_t i = 0; i < std::min(queries.size(), mFwd.size()); i++)
            regrow |= mFwd[i].dev.count < queries[i].nbDesc || mBwd[i].dev.count < train.nbDesc ||
                      mFwd[i].host.count < queries[i].nbDesc || mBwd[i].host.count < train.nbDesc;
        if (!regrow) return;
        hipCheck(hipStreamSynchronize(mStream), "reserveStorage");
        if (mFwd.size() < queries.size()) {
            mFwd.resize(queries.size());
            mBwd.resize(queries.size());
        }
        for (size_t i = 0; i < queries.size(); i++) {
            grow(mFwd[i].dev, queries[i].nbDesc);
            grow(mFwd[i].host, queries[i].nbDesc);
            grow(mBwd[i].dev, train.nbDesc);
            grow(mBwd[i].host, train.nbDesc);
        }
    }

    void matchImpl(const std::vector<Img>& queries, const Img& train)
    {
        reserveStorage(queries, train);
        if (mCfg.useCascadeHash) {
            std::vector<cashash::ImageDesc> qi(queries.size());
            std::vector<uint32_t> qs(queries.size());
            std::vector<cashash::MatchResult> res(queries.size());
            const PreComp& tp = mPreComp.at(train.hImage);
            for (size_t i = 0; i < queries.size(); i++) {
                const PreComp& p = mPreComp.at(queries[i].hImage);
                qi[i] = cashash::ImageDesc{queries[i].desc, p.ends.ptr, p.buckets.ptr, nullptr};
                qs[i] = queries[i].nbDesc;
                res[i] = cashash::MatchResult{mFwd[i].dev.ptr, mBwd[i].dev.ptr};
            }
            const cashash::ImageDesc ti{train.desc, tp.ends.ptr, tp.buckets.ptr, nullptr};
            cashash::findBestMatch(qi.data(), qs.data(), ti, train.nbDesc, res.data(), uint32_t(queries.size()), nullptr, 0,
                                   mStream);
        } else if (mCfg.useSift4b) {
            std::vector<Sift4bBruteForceMatchTask> tasks(queries.size());
            const PreComp& tp = mPreComp.at(train.hImage);
            for (size_t i = 0; i < queries.size(); i++) {
                const PreComp& p = mPreComp.at(queries[i].hImage);
                tasks[i].nbDescA = queries[i].nbDesc;
                tasks[i].nbDescB = train.nbDesc;
                tasks[i].descA = p.desc4.ptr;
                tasks[i].descB = tp.desc4.ptr;
                tasks[i].bestMatchA = mFwd[i].dev.ptr;
                tasks[i].bestMatchB = mBwd[i].dev.ptr;
                tasks[i].sqrNormDescA = nullptr;
                tasks[i].sqrNormDescB = nullptr;
            }
            launchBruteForceMatchIMMA(tasks.data(), tasks.size(), mStream);
        } else {
            std::vector<SiftBruteForceMatchTask> tasks(queries.size());
            for (size_t i = 0; i < queries.size(); i++)
                tasks[i] = SiftBruteForceMatchTask{queries[i].nbDesc, train.nbDesc, queries[i].desc, train.desc,
                                                   mFwd[i].dev.ptr, mBwd[i].dev.ptr, nullptr, nullptr};
            launchBruteForceMatchIMMA(tasks.data(), tasks.size(), mStream);
        }
    }

    Config mCfg;
    hipStream_t mStream = nullptr;
    std::mutex mLock;
    std::unordered_map<ImageHandle, PreComp> mPreComp;
    std::vector<MatchStorage> mFwd, mBwd;
    std::vector<KeptStorage> mKept;
    DeviceBuffer<uint32_t> mCounts;
    PinnedBuffer<uint32_t> mHostCounts;
    DeviceBuffer<int8_t> mWeights;
    DeviceBuffer<int32_t> mBiasBase;
    DeviceBuffer<uint64_t> mAcc;
    DeviceBuffer<int32_t> mBias;
    DeviceBuffer<uint8_t> mCodes;
    DeviceBuffer<uint32_t> mHist, mLower;
    uint64_t mAccNbDesc = 0;
    size_t mAllocations = 0;
};

}  // namespace rsfm
