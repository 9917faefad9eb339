/*
 * cashash.c - cascade hashing restated on the CPU.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * Restates (reference file:line):
 *  - kernelAccumulateDesc   casHash/accumulateDesc.cu:33-115  acc[k] += byte_k >> 1
 *  - kernelComputeBias      casHash/accumulateDesc.cu:136-174 FP32: negCenter = -float(acc)/float(n);
 *    8 lanes x 16 terms accumulated with fused multiply-add in k order
 *    (lane L owns k = 16L..16L+15), xor tree 4,2,1, roundf, + biasBase.
 *    PARITY UNPINNED against the reference: its fast-math division and FMA
 *    contraction are compiler-dependent (the reference's own test accepts
 *    +-1, test/casHashAccumulateDescTest.cpp:188); pinned here only to a
 *    float64 evaluation within +-1 and exactly on integer centres.
 *  - kernelComputeHash      casHash/computeHash.cu:437-472   bit c = bias[c] + sum_k w[c][k]*(byte_k>>1) >= 0,
 *    code byte s = channels 8s..8s+7 (bit b = channel 8s+b), remap word r bit b = channel 64+32r+b
 *    (bit mapping of test/casHashComputeHashTest.cpp:132-138).
 *  - kernelBuildHistogram / kernelComputeBucketBounds / kernelBuildBuckets
 *    casHash/buildBuckets.cu:29-260 (bucket order canonicalised: ascending index).
 *  - kernelFindBestMatch    casHash/findBestMatch.cu:570-617: exact L2^2 over the
 *    union of same-bucket candidates across schemes, lexicographic (dist, index) min.
 * New: six-table + 128-bit Hamming re-rank mode, seeded Gaussian weights.
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

void or_cashash_accumulate(const uint8_t* desc, uint32_t n, uint64_t acc[128])
{
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 128; k++) acc[k] += (uint64_t)(desc[(size_t)i * 128 + k] >> 1);
}

void or_cashash_compute_bias(const int8_t* w, const int32_t* bias_base, const uint64_t acc[128],
                             uint64_t nb_desc, uint32_t nch, int32_t* bias)
{
    float negc[128];
    for (int k = 0; k < 128; k++) negc[k] = -(float)acc[k] / (float)nb_desc;
    for (uint32_t c = 0; c < nch; c++) {
        float lane[8];
        for (int L = 0; L < 8; L++) {
            float v = 0.f;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    const int k = L * 16 + i * 4 + j;
                    v = fmaf(negc[k], (float)w[c * 128 + k], v);
                }
            lane[L] = v;
        }
        for (int m = 4; m >= 1; m /= 2) {
            float nl[8];
            for (int L = 0; L < 8; L++) nl[L] = lane[L] + lane[L ^ m];
            memcpy(lane, nl, sizeof(nl));
        }
        bias[c] = (int32_t)roundf(lane[0]) + bias_base[c];
    }
}

void or_cashash_compute_hash(const int8_t* w, const int32_t* bias, uint32_t nch, const uint8_t* desc,
                             uint32_t n, uint8_t* codes, uint32_t* remap)
{
    for (uint32_t i = 0; i < n; i++) {
        const uint8_t* d = desc + (size_t)i * 128;
        uint8_t code[8] = {0};
        uint32_t rm[4] = {0, 0, 0, 0};
        for (uint32_t c = 0; c < nch; c++) {
            int32_t s = bias[c];
            for (int k = 0; k < 128; k++) s += (int32_t)w[(size_t)c * 128 + k] * (int32_t)(d[k] >> 1);
            const int bit = s >= 0;
            if (c < 64) code[c / 8] |= (uint8_t)(bit << (c % 8));
            else rm[(c - 64) / 32] |= (uint32_t)bit << ((c - 64) % 32);
        }
        memcpy(codes + (size_t)i * 8, code, 8);
        if (remap && nch > 64) memcpy(remap + (size_t)i * 4, rm, 16);
    }
}

void or_cashash_histogram(const uint8_t* codes, uint32_t n, uint32_t hist[8 * 256])
{
    memset(hist, 0, sizeof(uint32_t) * 8 * 256);
    for (uint32_t i = 0; i < n; i++)
        for (int s = 0; s < 8; s++) hist[s * 256 + codes[(size_t)i * 8 + s]]++;
}

uint32_t or_cashash_bucket_bounds(const uint32_t hist[8 * 256], uint32_t lower[8 * 256],
                                  uint32_t upper[8 * 256])
{
    uint32_t mx = 0;
    for (int s = 0; s < 8; s++) {
        uint32_t sum = 0;
        for (int b = 0; b < 256; b++) {
            const uint32_t h = hist[s * 256 + b];
            lower[s * 256 + b] = sum;
            sum += h;
            upper[s * 256 + b] = sum;
            if (h > mx) mx = h;
        }
    }
    return mx;
}

void or_cashash_build_buckets(const uint8_t* codes, uint32_t n, const uint32_t lower[8 * 256],
                              uint32_t* buckets)
{
    uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * 8 * 256);
    memcpy(cur, lower, sizeof(uint32_t) * 8 * 256);
    for (uint32_t i = 0; i < n; i++)
        for (int s = 0; s < 8; s++) {
            const uint32_t b = codes[(size_t)i * 8 + s];
            buckets[(size_t)s * n + cur[s * 256 + b]++] = i;
        }
    free(cur);
}

static inline uint32_t l2(const uint8_t* a, const uint8_t* b)
{
    int32_t s = 0;
    for (int k = 0; k < 128; k++) {
        const int32_t d = (int32_t)a[k] - (int32_t)b[k];
        s += d * d;
    }
    return (uint32_t)s;
}

/* per-image bucket CSR built on the fly: order[s][...] sorted by code then index */
typedef struct { uint32_t* start; uint32_t* items; } csr_t;

static csr_t build_csr(const uint8_t* codes, uint32_t n, uint32_t nscheme)
{
    csr_t c;
    c.start = (uint32_t*)calloc((size_t)nscheme * 257, sizeof(uint32_t));
    c.items = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nscheme * (n ? n : 1));
    for (uint32_t s = 0; s < nscheme; s++) {
        uint32_t* st = c.start + (size_t)s * 257;
        for (uint32_t i = 0; i < n; i++) st[codes[(size_t)i * 8 + s] + 1]++;
        for (int b = 0; b < 256; b++) st[b + 1] += st[b];
        uint32_t cur[256];
        memcpy(cur, st, sizeof(cur));
        for (uint32_t i = 0; i < n; i++) c.items[(size_t)s * n + cur[codes[(size_t)i * 8 + s]]++] = i;
    }
    return c;
}

static void find_dir(const uint8_t* qdesc, const uint8_t* qcodes, uint32_t nq, const uint8_t* tdesc,
                     uint32_t nt, const csr_t* tc, uint32_t nscheme, or_best_match* best)
{
#pragma omp parallel for schedule(dynamic, 32)
    for (int64_t i = 0; i < (int64_t)nq; i++) {
        uint32_t bd = UINT32_MAX, bi = ~0u;
        for (uint32_t s = 0; s < nscheme; s++) {
            const uint32_t code = qcodes[(size_t)i * 8 + s];
            const uint32_t* st = tc->start + (size_t)s * 257;
            for (uint32_t p = st[code]; p < st[code + 1]; p++) {
                const uint32_t j = tc->items[(size_t)s * nt + p];
                const uint32_t d = l2(qdesc + (size_t)i * 128, tdesc + (size_t)j * 128);
                if (d < bd || (d == bd && j < bi)) { bd = d; bi = j; }
            }
        }
        best[i].index = bi;
        best[i].distance = bd;
    }
}

void or_cashash_find_best(const uint8_t* qdesc, const uint8_t* qcodes, uint32_t nq,
                          const uint8_t* tdesc, const uint8_t* tcodes, uint32_t nt,
                          uint32_t nscheme, or_best_match* bestQ, or_best_match* bestT)
{
    csr_t tc = build_csr(tcodes, nt, nscheme);
    find_dir(qdesc, qcodes, nq, tdesc, nt, &tc, nscheme, bestQ);
    free(tc.start); free(tc.items);
    if (bestT) {
        csr_t qc = build_csr(qcodes, nq, nscheme);
        find_dir(tdesc, tcodes, nt, qdesc, nq, &qc, nscheme, bestT);
        free(qc.start); free(qc.items);
    }
}

void or_cashash_find_best_rerank(const uint8_t* qdesc, const uint8_t* qcodes, const uint32_t* qremap,
                                 uint32_t nq, const uint8_t* tdesc, const uint8_t* tcodes,
                                 const uint32_t* tremap, uint32_t nt, uint32_t nscheme, uint32_t keep,
                                 or_best_match* bestQ)
{
    csr_t tc = build_csr(tcodes, nt, nscheme);
#pragma omp parallel
    {
        uint8_t* seen = (uint8_t*)calloc(nt ? nt : 1, 1);
        uint32_t cap = 1024;
        uint64_t* cand = (uint64_t*)malloc(sizeof(uint64_t) * cap);
#pragma omp for schedule(dynamic, 32)
        for (int64_t i = 0; i < (int64_t)nq; i++) {
            uint32_t nc = 0;
            for (uint32_t s = 0; s < nscheme; s++) {
                const uint32_t code = qcodes[(size_t)i * 8 + s];
                const uint32_t* st = tc.start + (size_t)s * 257;
                for (uint32_t p = st[code]; p < st[code + 1]; p++) {
                    const uint32_t j = tc.items[(size_t)s * nt + p];
                    if (seen[j]) continue;
                    seen[j] = 1;
                    uint32_t h = 0;
                    for (int r = 0; r < 4; r++)
                        h += (uint32_t)__builtin_popcount(qremap[(size_t)i * 4 + r] ^ tremap[(size_t)j * 4 + r]);
                    if (nc == cap) { cap *= 2; cand = (uint64_t*)realloc(cand, sizeof(uint64_t) * cap); }
                    cand[nc++] = ((uint64_t)h << 32) | j;
                }
            }
            for (uint32_t c = 0; c < nc; c++) seen[(uint32_t)cand[c]] = 0;
            /* partial selection of the `keep` least (hamming, index) keys */
            const uint32_t kk = keep < nc ? keep : nc;
            for (uint32_t a = 0; a < kk; a++) {
                uint32_t m = a;
                for (uint32_t b = a + 1; b < nc; b++) if (cand[b] < cand[m]) m = b;
                const uint64_t t = cand[a]; cand[a] = cand[m]; cand[m] = t;
            }
            uint32_t bd = UINT32_MAX, bi = ~0u;
            for (uint32_t a = 0; a < kk; a++) {
                const uint32_t j = (uint32_t)cand[a];
                const uint32_t d = l2(qdesc + (size_t)i * 128, tdesc + (size_t)j * 128);
                if (d < bd || (d == bd && j < bi)) { bd = d; bi = j; }
            }
            bestQ[i].index = bi;
            bestQ[i].distance = bd;
        }
        free(cand);
        free(seen);
    }
    free(tc.start); free(tc.items);
}

void or_cashash_gaussian_weights(uint64_t seed, uint32_t nch, int8_t* w, int32_t* bias_base)
{
    or_mt64 rng;
    or_mt64_seed(&rng, seed);
    for (size_t i = 0; i < (size_t)nch * 128; i++) {
        double v = 32.0 * or_mt64_normal(&rng);
        if (v < -128.0) v = -128.0;
        if (v > 127.0) v = 127.0;
        w[i] = (int8_t)lround(v);
    }
    for (uint32_t c = 0; c < nch; c++) {
        double s = 0;
        for (int k = 0; k < 128; k++) {
            double r = 16.0 * or_mt64_normal(&rng);
            if (r < -64.0) r = -64.0;
            if (r > 64.0) r = 64.0;
            s += (double)w[(size_t)c * 128 + k] * r;
        }
        bias_base[c] = (int32_t)lround(s);
    }
}
