// Exercises the reference-shaped C++ adapters (include/rsfm_amd.hpp) on the GPU:
// launch functions, cross-check, BruteForceMatcher (default, 4-bit and
// cascade-hash methods), RansacMatchFilter and error reporting. Prints one
// "PASS <name>" / "FAIL <name>" line per check; exit status 0 iff all pass.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>

#include "rsfm_amd.hpp"

using namespace rsfm;

// Layout stand-ins for cudapp::KArray (the reference's extern/cuda-utils, absent
// here): a plain array wrapper with a .data member, as the reference's call
// sites use it (casHashWeights.data, bias->data).
namespace cudapp {
template <typename T, uint32_t... Dims>
struct KArray;
template <typename T, uint32_t D0>
struct KArray<T, D0> {
    T data[D0];
};
template <typename T, uint32_t D0, uint32_t D1>
struct KArray<T, D0, D1> {
    KArray<T, D1> data[D0];
};
}  // namespace cudapp

static int g_fail = 0;
static void report(bool ok, const char* name)
{
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
    std::fflush(stdout);
    if (!ok) g_fail++;
}

static std::vector<uint8_t> random_desc(std::mt19937& rng, uint32_t n)
{
    std::vector<uint8_t> d(size_t(n) * 128);
    for (auto& x : d) x = uint8_t(rng() & 0xFF);
    return d;
}

static uint32_t l2(const uint8_t* a, const uint8_t* b, int dims, bool nibbleHigh)
{
    uint32_t s = 0;
    for (int k = 0; k < dims; k++) {
        const int x = nibbleHigh ? (a[k] >> 4) : a[k], y = nibbleHigh ? (b[k] >> 4) : b[k];
        s += uint32_t((x - y) * (x - y));
    }
    return s;
}

// exact top-1 both ways, lowest index on ties
static void cpu_match(const std::vector<uint8_t>& A, uint32_t nA, const std::vector<uint8_t>& B, uint32_t nB,
                      std::vector<BestMatch>& fa, std::vector<BestMatch>& fb, bool high = false)
{
    fa.assign(nA, BestMatch{~0u, ~0u});
    fb.assign(nB, BestMatch{~0u, ~0u});
    for (uint32_t i = 0; i < nA; i++)
        for (uint32_t j = 0; j < nB; j++) {
            const uint32_t d = l2(&A[size_t(i) * 128], &B[size_t(j) * 128], 128, high);
            if (d < fa[i].distance) fa[i] = BestMatch{j, d};
            if (d < fb[j].distance) fb[j] = BestMatch{i, d};
        }
}

static bool same(const std::vector<BestMatch>& a, const std::vector<BestMatch>& b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (a[i].index != b[i].index || a[i].distance != b[i].distance) return false;
    return true;
}

template <typename T>
static T* upload(const std::vector<T>& h)
{
    T* d = nullptr;
    hipCheck(hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)), "hipMalloc");
    hipCheck(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "H2D");
    return d;
}

int main()
{
    std::mt19937 rng(7);
    const uint32_t nA = 700, nB = 900;
    auto A = random_desc(rng, nA), B = random_desc(rng, nB);
    // plant duplicates so ties occur
    std::memcpy(&B[5 * 128], &A[3 * 128], 128);
    std::memcpy(&B[600 * 128], &A[3 * 128], 128);
    std::vector<BestMatch> ra, rb;
    cpu_match(A, nA, B, nB, ra, rb);
    uint8_t* dA = upload(A);
    uint8_t* dB = upload(B);
    hipStream_t s;
    hipCheck(hipStreamCreate(&s), "stream");

    // launch functions
    {
        DeviceBuffer<uint32_t> na(nA);
        launchPreCompSiftSqrNorm(na.ptr, dA, nA, s);
        std::vector<uint32_t> h(nA);
        hipCheck(hipMemcpyAsync(h.data(), na.ptr, nA * 4, hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipStreamSynchronize(s), "sync");
        bool ok = true;
        for (uint32_t i = 0; i < nA; i++) ok &= h[i] == l2(&A[size_t(i) * 128], std::vector<uint8_t>(128, 0).data(), 128, false);
        report(ok, "launchPreCompSiftSqrNorm");
        for (int impl = 0; impl < 2; impl++) {
            DeviceBuffer<BestMatch> ba(nA), bb(nB);
            SiftBruteForceMatchTask t{nA, nB, dA, dB, ba.ptr, bb.ptr, na.ptr, nullptr};
            if (impl == 0) launchBruteForceMatchIMMA(&t, 1, s);
            else launchBruteForceMatchIDP(&t, 1, s);
            std::vector<BestMatch> ha(nA), hb(nB);
            hipCheck(hipMemcpyAsync(ha.data(), ba.ptr, nA * 8, hipMemcpyDeviceToHost, s), "D2H");
            hipCheck(hipMemcpyAsync(hb.data(), bb.ptr, nB * 8, hipMemcpyDeviceToHost, s), "D2H");
            hipCheck(hipStreamSynchronize(s), "sync");
            report(same(ha, ra) && same(hb, rb), impl == 0 ? "launchBruteForceMatchIMMA" : "launchBruteForceMatchIDP");
        }
    }
    // cross-check on the host (bruteForceMatch.cpp:23-40)
    {
        auto v = crossCheckMatches(ra.data(), nA, rb.data(), nB);
        bool ok = !v.empty();
        for (const auto& m : v) ok &= rb[m.idxTrain].index == m.idxQuery && ra[m.idxQuery].index == m.idxTrain;
        size_t expect = 0;
        for (uint32_t i = 0; i < nA; i++) expect += rb[ra[i].index].index == i;
        report(ok && v.size() == expect, "crossCheckMatches");
        auto r = removeMatchConflicts(ra.data(), nA, nB);
        report(!r.empty() && r.size() <= nB, "removeMatchConflicts");
    }
    // BruteForceMatcher: default method, two queries against one train
    {
        BruteForceMatcher m;
        auto A2 = random_desc(rng, 333);
        uint8_t* dA2 = upload(A2);
        using D = BruteForceMatcher::Descriptor;
        m.registerImage(0, (const D*)dA, nA, s);
        m.registerImage(1, (const D*)dA2, 333, s);
        m.registerImage(2, (const D*)dB, nB, s);
        auto res = m.match({{(const D*)dA, nA, 0}, {(const D*)dA2, 333, 1}}, {(const D*)dB, nB, 2}, true, s).get();
        std::vector<BestMatch> qa, qb;
        cpu_match(A2, 333, B, nB, qa, qb);
        report(res.size() == 2 && same(res[0].first, ra) && same(res[0].second, rb) && same(res[1].first, qa) &&
                   same(res[1].second, qb),
               "BruteForceMatcher::match");
        auto filt = m.matchAndFilter({{(const D*)dA, nA, 0}}, {(const D*)dB, nB, 2}, true, s).get();
        auto cc = crossCheckMatches(ra.data(), nA, rb.data(), nB);
        bool ok = filt.size() == 1 && filt[0].size() == cc.size();
        for (size_t k = 0; ok && k < cc.size(); k++)
            ok &= filt[0][k].idxQuery == cc[k].idxQuery && filt[0][k].idxTrain == cc[k].idxTrain &&
                  filt[0][k].distance == cc[k].distance;
        report(ok, "BruteForceMatcher::matchAndFilter (device cross-check)");
        auto filt2 = m.matchAndFilter({{(const D*)dA, nA, 0}, {(const D*)dA2, 333, 1}}, {(const D*)dB, nB, 2}, false, s).get();
        auto rc = removeMatchConflicts(ra.data(), nA, nB);
        bool ok2 = filt2.size() == 2 && filt2[0].size() == rc.size();
        for (size_t k = 0; ok2 && k < rc.size(); k++)
            ok2 &= filt2[0][k].idxQuery == rc[k].idxQuery && filt2[0][k].idxTrain == rc[k].idxTrain;
        report(ok2, "BruteForceMatcher::matchAndFilter (conflicts)");
        hipCheck(hipFree(dA2), "free");
    }
    // GPU-resident service: after a warm-up call the storage is reused, no
    // allocation happens over ten further calls (including smaller ones), and
    // every result is right
    {
        BruteForceMatcher m;
        using D = BruteForceMatcher::Descriptor;
        m.registerImage(0, (const D*)dA, nA, s);
        m.registerImage(2, (const D*)dB, nB, s);
        auto cc = crossCheckMatches(ra.data(), nA, rb.data(), nB);
        (void)m.match({{(const D*)dA, nA, 0}, {(const D*)dA, nA, 0}}, {(const D*)dB, nB, 2}, true, s).get();
        (void)m.matchAndFilter({{(const D*)dA, nA, 0}, {(const D*)dA, nA, 0}}, {(const D*)dB, nB, 2}, true, s).get();
        const size_t warm = m.allocations();
        bool ok = warm > 0;
        std::vector<std::future<BruteForceMatcher::Result>> pending;
        for (int it = 0; it < 10; it++) {
            if (it % 2 == 0) {
                pending.push_back(m.match({{(const D*)dA, nA, 0}}, {(const D*)dB, nB, 2}, true, s));
            } else {
                auto f = m.matchAndFilter({{(const D*)dA, nA, 0}, {(const D*)dA, 300, 0}}, {(const D*)dB, nB, 2}, true, s)
                             .get();
                ok &= f.size() == 2 && f[0].size() == cc.size();
                for (size_t k = 0; ok && k < cc.size(); k++)
                    ok &= f[0][k].idxQuery == cc[k].idxQuery && f[0][k].idxTrain == cc[k].idxTrain;
            }
        }
        for (auto& f : pending) {
            auto r = f.get();
            ok &= r.size() == 1 && same(r[0].first, ra) && same(r[0].second, rb);
        }
        ok &= m.allocations() == warm;
        report(ok, "BruteForceMatcher storage reused over 10 calls (no allocation after warm-up)");
    }
    // the reference's cascade-hash registration sequence, written as at
    // bruteForceMatcher.cpp:97-103 against KArray-typed device objects, gives
    // the same codes, bounds and bucket sets as the raw-pointer entry points
    {
        using cudapp::KArray;
        std::vector<int8_t> w(64 * 128);
        std::vector<int32_t> base(64);
        std::normal_distribution<float> nd(0.f, 32.f);
        for (auto& x : w) x = int8_t(std::max(-128.f, std::min(127.f, std::round(nd(rng)))));
        for (auto& x : base) x = int32_t(rng() % 2001) - 1000;
        auto* casHashWeightsPtr = reinterpret_cast<KArray<int32_t, 64, 32>*>(upload(w));
        auto* casHashBiasBasePtr = reinterpret_cast<KArray<int32_t, 64>*>(upload(base));
        const KArray<int32_t, 64, 32>& casHashWeights = *casHashWeightsPtr;
        const KArray<int32_t, 64>& casHashBiasBase = *casHashBiasBasePtr;
        KArray<uint64_t, 128>* accPtr = nullptr;
        KArray<int32_t, 64>* bias = nullptr;
        KArray<uint8_t, cashash::nbScheme>* bucketIndices = nullptr;
        KArray<uint32_t, cashash::nbScheme, cashash::nbBuckets>* histAndBounds = nullptr;
        KArray<uint32_t, cashash::nbScheme, cashash::nbBuckets>* casHashBucketEnds = nullptr;
        uint32_t* casHashBuckets = nullptr;
        uint32_t* pinnedMaxBucketSize = nullptr;
        hipCheck(hipMalloc(&accPtr, sizeof(*accPtr)), "malloc");
        hipCheck(hipMalloc(&bias, sizeof(*bias)), "malloc");
        hipCheck(hipMalloc(&bucketIndices, sizeof(*bucketIndices) * nB), "malloc");
        hipCheck(hipMalloc(&histAndBounds, sizeof(*histAndBounds) * 2), "malloc");
        hipCheck(hipMalloc(&casHashBucketEnds, sizeof(*casHashBucketEnds)), "malloc");
        hipCheck(hipMalloc(&casHashBuckets, sizeof(uint32_t) * nB * cashash::nbScheme), "malloc");
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&pinnedMaxBucketSize), 4, hipHostMallocDefault), "pinned");
        KArray<uint64_t, 128>& acc = *accPtr;
        auto* const histogram = histAndBounds;
        auto* const idxAllocator = histAndBounds + 1;
        const auto* src = reinterpret_cast<const KArray<uint32_t, 32>*>(dB);
        const uint32_t nbDesc = nB;
        uint64_t accNbDesc = 0;
        const bool isFirst = accNbDesc == 0;
        hipStream_t stream = s;
        // the reference's lines, unchanged apart from the stream type
        cashash::accumuateDesc(acc, nullptr, isFirst, src, nbDesc, stream);
        accNbDesc += nbDesc;
        cashash::computeBias(bias[0], casHashWeights, casHashBiasBase, acc, accNbDesc, stream);
        cashash::computeHashForBucketOnly(casHashWeights.data, bias->data, src, nbDesc, bucketIndices, nullptr, stream);
        cashash::buildHistogram(histogram, bucketIndices, nbDesc, stream);
        cashash::computeBucketBound(*idxAllocator, *casHashBucketEnds, pinnedMaxBucketSize, *histogram, stream);
        cashash::buildBuckets(bucketIndices, nbDesc, *idxAllocator, casHashBuckets, stream);
        // the same through the raw-pointer entry points
        DeviceBuffer<uint64_t> acc2(128);
        DeviceBuffer<int32_t> bias2(64);
        DeviceBuffer<uint8_t> codes2(size_t(nB) * 8);
        DeviceBuffer<uint32_t> hist2(2048), lower2(2048), ends2(2048), buckets2(size_t(nB) * 8);
        rsfmCheck(rsfm_cashash_accumulate_desc(acc2.ptr, nullptr, 1, dB, nB, s), "acc");
        rsfmCheck(rsfm_cashash_compute_bias(bias2.ptr, reinterpret_cast<const int8_t*>(casHashWeightsPtr),
                                            reinterpret_cast<const int32_t*>(casHashBiasBasePtr), acc2.ptr, nB, 64, s),
                  "bias");
        rsfmCheck(rsfm_cashash_compute_hash_bucket_only(reinterpret_cast<const int8_t*>(casHashWeightsPtr), bias2.ptr, dB,
                                                        nB, codes2.ptr, nullptr, s),
                  "hash");
        rsfmCheck(rsfm_cashash_build_histogram(hist2.ptr, codes2.ptr, nB, s), "hist");
        rsfmCheck(rsfm_cashash_compute_bucket_bound(lower2.ptr, ends2.ptr, nullptr, hist2.ptr, s), "bounds");
        rsfmCheck(rsfm_cashash_build_buckets(codes2.ptr, nB, lower2.ptr, buckets2.ptr, s), "buckets");
        std::vector<uint8_t> c1(size_t(nB) * 8), c2(size_t(nB) * 8);
        std::vector<uint32_t> e1(2048), e2(2048), b1(size_t(nB) * 8), b2(size_t(nB) * 8);
        hipCheck(hipMemcpyAsync(c1.data(), bucketIndices, c1.size(), hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipMemcpyAsync(c2.data(), codes2.ptr, c2.size(), hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipMemcpyAsync(e1.data(), casHashBucketEnds, 8192, hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipMemcpyAsync(e2.data(), ends2.ptr, 8192, hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipMemcpyAsync(b1.data(), casHashBuckets, b1.size() * 4, hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipMemcpyAsync(b2.data(), buckets2.ptr, b2.size() * 4, hipMemcpyDeviceToHost, s), "D2H");
        hipCheck(hipStreamSynchronize(s), "sync");
        bool ok = c1 == c2 && e1 == e2 && *pinnedMaxBucketSize > 0;
        for (uint32_t sc = 0; ok && sc < 8; sc++)
            for (uint32_t bk = 0; bk < 256; bk++) {
                const uint32_t lo = bk ? e1[sc * 256 + bk - 1] : 0, hi = e1[sc * 256 + bk];
                std::vector<uint32_t> x(b1.begin() + sc * nB + lo, b1.begin() + sc * nB + hi);
                std::vector<uint32_t> y(b2.begin() + sc * nB + lo, b2.begin() + sc * nB + hi);
                std::sort(x.begin(), x.end());
                std::sort(y.begin(), y.end());
                ok &= x == y;
            }
        report(ok, "cashash call sites with KArray-typed arguments (bruteForceMatcher.cpp:97-103)");
        for (void* p : {(void*)casHashWeightsPtr, (void*)casHashBiasBasePtr, (void*)accPtr, (void*)bias,
                        (void*)bucketIndices, (void*)histAndBounds, (void*)casHashBucketEnds, (void*)casHashBuckets})
            hipCheck(hipFree(p), "free");
        hipCheck(hipHostFree(pinnedMaxBucketSize), "free");
    }
    // 4-bit method: L2 on the high nibbles
    {
        BruteForceMatcher::Config cfg;
        cfg.useSift4b = true;
        BruteForceMatcher m(cfg);
        using D = BruteForceMatcher::Descriptor;
        m.registerImage(0, (const D*)dA, nA, s);
        m.registerImage(2, (const D*)dB, nB, s);
        auto res = m.match({{(const D*)dA, nA, 0}}, {(const D*)dB, nB, 2}, true, s).get();
        std::vector<BestMatch> fa, fb;
        cpu_match(A, nA, B, nB, fa, fb, true);
        report(same(res[0].first, fa) && same(res[0].second, fb), "BruteForceMatcher(useSift4b)::match");
    }
    // cascade-hash method: a query identical to the train image matches itself
    {
        BruteForceMatcher::Config cfg;
        cfg.useCascadeHash = true;
        cfg.casHashWeights.resize(64 * 128);
        cfg.casHashBiasBase.assign(64, 0);
        std::normal_distribution<float> nd(0.f, 32.f);
        for (auto& w : cfg.casHashWeights) w = int8_t(std::max(-128.f, std::min(127.f, std::round(nd(rng)))));
        BruteForceMatcher m(cfg);
        using D = BruteForceMatcher::Descriptor;
        m.registerImage(0, (const D*)dB, nB, s);
        m.registerImage(1, (const D*)dB, nB, s);
        auto res = m.match({{(const D*)dB, nB, 0}}, {(const D*)dB, nB, 1}, true, s).get();
        // image 1 was hashed with the mean of both registrations (same data), so
        // its codes equal image 0's and every descriptor finds itself (or an
        // identical copy at a lower index) at distance 0
        bool ok = true;
        for (uint32_t i = 0; i < nB; i++) {
            const auto& f = res[0].first[i];
            ok &= f.distance == 0 && f.index <= i && l2(&B[size_t(i) * 128], &B[size_t(f.index) * 128], 128, false) == 0;
        }
        report(ok, "BruteForceMatcher(useCascadeHash)::match");
    }
    // RansacMatchFilter on an affine scene with outliers
    {
        RansacMatchFilter f;
        std::uniform_real_distribution<float> ux(0.f, 4000.f), uy(0.f, 3000.f);
        std::vector<std::pair<Vec2f, Vec2f>> mm;
        std::vector<bool> truth;
        for (int k = 0; k < 3000; k++) {
            const float x = ux(rng), y = uy(rng);
            const bool in = k % 4 != 0;
            if (in) mm.push_back({{x, y}, {1.02f * x + 0.05f * y + 10.f, -0.03f * x + 0.99f * y - 5.f}});
            else mm.push_back({{x, y}, {ux(rng), uy(rng)}});
            truth.push_back(in);
        }
        auto mask = f.getInlierMask({4000, 3000}, mm, 2, 256, 0.15f, 16, true, s, 42);
        size_t tp = 0, fp = 0, nin = 0;
        for (size_t k = 0; k < mm.size(); k++) {
            nin += truth[k];
            tp += truth[k] && mask[k];
            fp += !truth[k] && mask[k];
        }
        report(tp > 0.95 * nin && fp < 0.02 * (mm.size() - nin), "RansacMatchFilter::getInlierMask");
    }
    // errors surface as exceptions with the library message
    {
        bool threw = false;
        try {
            SiftBruteForceMatchTask t{5, 5, nullptr, dB, nullptr, nullptr, nullptr, nullptr};
            launchBruteForceMatchIMMA(&t, 1, s);
        } catch (const Exception& e) {
            threw = std::strstr(e.what(), "descA") != nullptr;
        }
        report(threw, "error reporting");
    }
    hipCheck(hipStreamDestroy(s), "destroy");
    hipCheck(hipFree(dA), "free");
    hipCheck(hipFree(dB), "free");
    std::printf("%s\n", g_fail ? "SOME FAILED" : "ALL PASSED");
    return g_fail ? 1 : 0;
}
