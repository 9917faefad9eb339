/*
 * mlp.c - forward/backward of the learned RANSAC proposal network.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * No reference counterpart: the reference's only biased sampler is
 * NonDuplicateSampler::Picker's focused tail (ransac.hpp:28-45). The network
 * (8 -> 64 -> 64 -> 64 -> 1, ReLU, sigmoid, mean binary cross-entropy) is this
 * project's own.
 *
 * Two modes share one code path:
 *  - exact (q = 0): everything in FP64. Pinned in tests by central finite
 *    differences of the loss and by closed forms at zero weights.
 *  - bf16 emulation (q = 1): the same arithmetic with the roundings of the GPU
 *    kernels inserted: weights and inputs rounded to bf16 (round to nearest
 *    even), biases to FP32, every hidden activation and every back-propagated
 *    delta rounded to bf16 where the kernel feeds it to a bf16 matrix
 *    instruction, dz rounded to bf16 for the output-layer weight gradient.
 *    Sums stay in FP64 (the kernels accumulate in FP32). With q = 0 the
 *    rounding calls are identities, so this mode differs from the pinned one
 *    only by those calls.
 */
#include "rsfm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define H 64
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + H * 8)
#define OFF_W2 (OFF_B1 + H)
#define OFF_B2 (OFF_W2 + H * H)
#define OFF_W3 (OFF_B2 + H)
#define OFF_B3 (OFF_W3 + H * H)
#define OFF_W4 (OFF_B3 + H)
#define OFF_B4 (OFF_W4 + H)

/* float -> bf16 -> float, round to nearest even (NaN kept quiet) */
double or_bf16_round(double v)
{
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
        u |= 0x00400000u;
    } else {
        u += 0x7FFFu + ((u >> 16) & 1u);
    }
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return f;
}

typedef struct {
    int q;
    double w1[H * 8], b1[H], w2[H * H], b2[H], w3[H * H], b3[H], w4[H], b4;
} Net;

static double rq(int q, double v) { return q ? or_bf16_round(v) : v; }
static double rf(int q, double v) { return q ? (double)(float)v : v; }

static void load_net(Net* n, const double* p, int q)
{
    n->q = q;
    for (int i = 0; i < H * 8; i++) n->w1[i] = rq(q, p[OFF_W1 + i]);
    for (int i = 0; i < H * H; i++) n->w2[i] = rq(q, p[OFF_W2 + i]);
    for (int i = 0; i < H * H; i++) n->w3[i] = rq(q, p[OFF_W3 + i]);
    for (int i = 0; i < H; i++) {
        n->w4[i] = rq(q, p[OFF_W4 + i]);
        n->b1[i] = rf(q, p[OFF_B1 + i]);
        n->b2[i] = rf(q, p[OFF_B2 + i]);
        n->b3[i] = rf(q, p[OFF_B3 + i]);
    }
    n->b4 = rf(q, p[OFF_B4]);
}

/* forward of one sample; h1..h3 are the (rounded) activations fed onwards */
static double forward_one(const Net* n, const float* x, double xr[8], double h1[H], double h2[H], double h3[H])
{
    const int q = n->q;
    for (int i = 0; i < 8; i++) xr[i] = rq(q, x[i]);
    for (int o = 0; o < H; o++) {
        double s = n->b1[o];
        for (int i = 0; i < 8; i++) s += n->w1[o * 8 + i] * xr[i];
        s = rf(q, s);
        h1[o] = rq(q, s > 0 ? s : 0);
    }
    for (int o = 0; o < H; o++) {
        double s = n->b2[o];
        for (int i = 0; i < H; i++) s += n->w2[o * H + i] * h1[i];
        s = rf(q, s);
        h2[o] = rq(q, s > 0 ? s : 0);
    }
    for (int o = 0; o < H; o++) {
        double s = n->b3[o];
        for (int i = 0; i < H; i++) s += n->w3[o * H + i] * h2[i];
        s = rf(q, s);
        h3[o] = rq(q, s > 0 ? s : 0);
    }
    double z = n->b4;
    for (int i = 0; i < H; i++) z += n->w4[i] * h3[i];
    return rf(q, z);
}

static void forward_all(const double* params, const float* x, uint32_t cnt, double* prob, int q)
{
    Net* n = (Net*)malloc(sizeof(Net));
    load_net(n, params, q);
    double xr[8], h1[H], h2[H], h3[H];
    for (uint32_t k = 0; k < cnt; k++) {
        const double z = forward_one(n, x + (size_t)k * 8, xr, h1, h2, h3);
        prob[k] = 1.0 / (1.0 + exp(-z));
    }
    free(n);
}

static double loss_grad_all(const double* params, const float* x, const float* y, uint32_t cnt, double* grad, int q)
{
    if (grad) memset(grad, 0, sizeof(double) * OR_MLP_NPARAM);
    Net* n = (Net*)malloc(sizeof(Net));
    load_net(n, params, q);
    double loss = 0;
    double xr[8], h1[H], h2[H], h3[H], g1[H], g2[H], g3[H];
    for (uint32_t k = 0; k < cnt; k++) {
        const float* xi = x + (size_t)k * 8;
        const double z = forward_one(n, xi, xr, h1, h2, h3);
        const double yy = y[k];
        loss += (z > 0 ? z : 0) - yy * z + log1p(exp(-fabs(z)));
        if (!grad) continue;
        const double pz = 1.0 / (1.0 + exp(-z));
        const double dz = rf(q, (pz - yy) / cnt);
        const double dzq = rq(q, dz);
        grad[OFF_B4] += dzq;
        for (int i = 0; i < H; i++) {
            grad[OFF_W4 + i] += dzq * h3[i];
            g3[i] = rq(q, h3[i] > 0 ? dz * n->w4[i] : 0);
        }
        for (int o = 0; o < H; o++) {
            grad[OFF_B3 + o] += g3[o];
            for (int i = 0; i < H; i++) grad[OFF_W3 + o * H + i] += g3[o] * h2[i];
        }
        for (int i = 0; i < H; i++) {
            double s = 0;
            for (int o = 0; o < H; o++) s += n->w3[o * H + i] * g3[o];
            g2[i] = rq(q, h2[i] > 0 ? rf(q, s) : 0);
        }
        for (int o = 0; o < H; o++) {
            grad[OFF_B2 + o] += g2[o];
            for (int i = 0; i < H; i++) grad[OFF_W2 + o * H + i] += g2[o] * h1[i];
        }
        for (int i = 0; i < H; i++) {
            double s = 0;
            for (int o = 0; o < H; o++) s += n->w2[o * H + i] * g2[o];
            g1[i] = rq(q, h1[i] > 0 ? rf(q, s) : 0);
        }
        for (int o = 0; o < H; o++) {
            grad[OFF_B1 + o] += g1[o];
            for (int i = 0; i < 8; i++) grad[OFF_W1 + o * 8 + i] += g1[o] * xr[i];
        }
    }
    free(n);
    return loss / cnt;
}

void or_mlp_forward(const double* params, const float* x, uint32_t n, double* prob)
{
    forward_all(params, x, n, prob, 0);
}

double or_mlp_loss_grad(const double* params, const float* x, const float* y, uint32_t n, double* grad)
{
    return loss_grad_all(params, x, y, n, grad, 0);
}

void or_mlp_forward_bf16(const double* params, const float* x, uint32_t n, double* prob)
{
    forward_all(params, x, n, prob, 1);
}

double or_mlp_loss_grad_bf16(const double* params, const float* x, const float* y, uint32_t n, double* grad)
{
    return loss_grad_all(params, x, y, n, grad, 1);
}

/* Input features of the proposal network (rsfm_proposal_features_batch). FP32
 * like the kernel; the kernel's sqrtf/log2f may differ in the last ulp. */
void or_proposal_features(const or_valid_match* m, uint32_t n, const or_best_match* second, const float* kpQ,
                          const float* kpT, float width, float height, float* x)
{
    const float pi = 3.14159265358979f;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t q = m[k].idxQuery, t = m[k].idxTrain;
        const float d1 = m[k].distance;
        const uint32_t d2 = second[q].distance;
        float* f = x + (size_t)k * 8;
        if (second[q].index == 0xFFFFFFFFu)
            f[0] = 0.f;
        else if (d2 == 0)
            f[0] = 1.f;
        else {
            const float r = sqrtf(d1 / (float)d2);
            f[0] = r < 1.f ? r : 1.f;
        }
        f[1] = sqrtf(d1) * (1.f / 512.f);
        const float* a = kpQ + (size_t)q * 4;
        const float* b = kpT + (size_t)t * 4;
        f[2] = a[0] * (1.f / width) * 2.f - 1.f;
        f[3] = a[1] * (1.f / height) * 2.f - 1.f;
        f[4] = b[0] * (1.f / width) * 2.f - 1.f;
        f[5] = b[1] * (1.f / height) * 2.f - 1.f;
        float ls = log2f(b[2] / a[2]);
        f[6] = ls < -8.f ? -8.f : (ls > 8.f ? 8.f : ls);
        float d = b[3] - a[3];
        d = d - 2.f * pi * floorf((d + pi) * (0.5f / pi));
        f[7] = d * (1.f / pi);
    }
}
