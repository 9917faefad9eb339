/*
 * mlp.c - FP64 forward/backward of the learned RANSAC proposal network.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * No reference counterpart: the reference's only biased sampler is
 * NonDuplicateSampler::Picker's focused tail (ransac.hpp:28-45). The network
 * (8 -> 64 -> 64 -> 64 -> 1, ReLU, sigmoid, mean binary cross-entropy) is this
 * project's own. Pinned in tests by central finite differences of the FP64
 * loss and by closed forms at zero weights (p = sigmoid(b4)).
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

static double forward_one(const double* p, const float* x, double h1[H], double h2[H], double h3[H])
{
    for (int o = 0; o < H; o++) {
        double s = p[OFF_B1 + o];
        for (int i = 0; i < 8; i++) s += p[OFF_W1 + o * 8 + i] * (double)x[i];
        h1[o] = s > 0 ? s : 0;
    }
    for (int o = 0; o < H; o++) {
        double s = p[OFF_B2 + o];
        for (int i = 0; i < H; i++) s += p[OFF_W2 + o * H + i] * h1[i];
        h2[o] = s > 0 ? s : 0;
    }
    for (int o = 0; o < H; o++) {
        double s = p[OFF_B3 + o];
        for (int i = 0; i < H; i++) s += p[OFF_W3 + o * H + i] * h2[i];
        h3[o] = s > 0 ? s : 0;
    }
    double z = p[OFF_B4];
    for (int i = 0; i < H; i++) z += p[OFF_W4 + i] * h3[i];
    return z;
}

void or_mlp_forward(const double* params, const float* x, uint32_t n, double* prob)
{
    double h1[H], h2[H], h3[H];
    for (uint32_t k = 0; k < n; k++) {
        const double z = forward_one(params, x + (size_t)k * 8, h1, h2, h3);
        prob[k] = 1.0 / (1.0 + exp(-z));
    }
}

double or_mlp_loss_grad(const double* params, const float* x, const float* y, uint32_t n, double* grad)
{
    if (grad) memset(grad, 0, sizeof(double) * OR_MLP_NPARAM);
    double loss = 0;
    double h1[H], h2[H], h3[H], g1[H], g2[H], g3[H];
    for (uint32_t k = 0; k < n; k++) {
        const float* xi = x + (size_t)k * 8;
        const double z = forward_one(params, xi, h1, h2, h3);
        const double yy = y[k];
        loss += (z > 0 ? z : 0) - yy * z + log1p(exp(-fabs(z)));
        if (!grad) continue;
        const double pz = 1.0 / (1.0 + exp(-z));
        const double dz = (pz - yy) / n;
        grad[OFF_B4] += dz;
        for (int i = 0; i < H; i++) {
            grad[OFF_W4 + i] += dz * h3[i];
            g3[i] = h3[i] > 0 ? dz * params[OFF_W4 + i] : 0;
        }
        for (int o = 0; o < H; o++) {
            grad[OFF_B3 + o] += g3[o];
            for (int i = 0; i < H; i++) grad[OFF_W3 + o * H + i] += g3[o] * h2[i];
        }
        for (int i = 0; i < H; i++) {
            double s = 0;
            for (int o = 0; o < H; o++) s += params[OFF_W3 + o * H + i] * g3[o];
            g2[i] = h2[i] > 0 ? s : 0;
        }
        for (int o = 0; o < H; o++) {
            grad[OFF_B2 + o] += g2[o];
            for (int i = 0; i < H; i++) grad[OFF_W2 + o * H + i] += g2[o] * h1[i];
        }
        for (int i = 0; i < H; i++) {
            double s = 0;
            for (int o = 0; o < H; o++) s += params[OFF_W2 + o * H + i] * g2[o];
            g1[i] = h1[i] > 0 ? s : 0;
        }
        for (int o = 0; o < H; o++) {
            grad[OFF_B1 + o] += g1[o];
            for (int i = 0; i < 8; i++) grad[OFF_W1 + o * 8 + i] += g1[o] * (double)xi[i];
        }
    }
    return loss / n;
}
