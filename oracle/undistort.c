/*
 * undistort.c - keypoint undistortion restated on the CPU.
 * TEST INFRASTRUCTURE ONLY (see rsfm_oracle.h).
 *
 * Reference: undistort / DistortParams (distortion.cpp:63-210): k1, k2, p1,
 * p2, k3 model, Jacobian, Gauss-Newton on J^T J delta = J^T error; the
 * in-place driver (distortion.cpp:222-272) maps pixels to normalised
 * coordinates and back. Repeats rsfm_undistort_in_place expression for
 * expression (FP32, no contraction, Cramer 2x2 solve). Pinned in tests by
 * distortion of the result reproducing the input, an FP64 Newton solve run to
 * convergence, and the identity for zero distortion. The reference is built
 * with -ffast-math (distortion.cpp:17-19), so its own FP32 bits are not
 * reproducible.
 */
#include "rsfm_oracle.h"
#include <math.h>

static void value_jacobian(const float k[5], float x0, float x1, float* v0, float* v1, float J[2][2])
{
    const float k1 = k[0], k2 = k[1], p1 = k[2], p2 = k[3], k3 = k[4];
    const float r2 = x0 * x0 + x1 * x1;
    const float scale = ((1.f + k1 * r2) + k2 * (r2 * r2)) + k3 * (r2 * r2 * r2);
    *v0 = x0 * scale + ((2.f * p1 * (x0 * x1) + p2 * r2) + 2.f * p2 * (x0 * x0));
    *v1 = x1 * scale + ((2.f * p2 * (x0 * x1) + p1 * r2) + 2.f * p1 * (x1 * x1));
    const float sj = (k1 + 2.f * k2 * r2) + 3.f * k3 * (r2 * r2);
    const float j01 = ((2.f * x0 * x1 * sj) + 2.f * p1 * x0) + 2.f * p2 * x1;
    J[0][0] = ((scale + 2.f * sj * (x0 * x0)) + 2.f * p1 * x1) + 6.f * p2 * x0;
    J[0][1] = j01;
    J[1][0] = j01;
    J[1][1] = ((scale + 2.f * sj * (x1 * x1)) + 6.f * p1 * x1) + 2.f * p2 * x0;
}

void or_distort_f32(const float k[5], float x0, float x1, float out[2])
{
    float J[2][2];
    value_jacobian(k, x0, x1, &out[0], &out[1], J);
}

void or_undistort_in_place(float* px, float* py, uint32_t n, const float cam[4], const float k[5], uint32_t iters)
{
    if (k[0] == 0.f && k[1] == 0.f && k[2] == 0.f && k[3] == 0.f && k[4] == 0.f) return;
    const float ifx = 1.f / cam[0], ify = 1.f / cam[1];
    const float icx = -cam[2] * ifx, icy = -cam[3] * ify;
    for (uint32_t i = 0; i < n; i++) {
        const float tx = px[i] * ifx + icx, ty = py[i] * ify + icy;
        float sx = tx, sy = ty;
        for (uint32_t it = 0; it < iters; it++) {
            float v0, v1, J[2][2];
            value_jacobian(k, sx, sy, &v0, &v1, J);
            const float e0 = tx - v0, e1 = ty - v1;
            const float h00 = J[0][0] * J[0][0] + J[1][0] * J[1][0];
            const float h01 = J[0][0] * J[0][1] + J[1][0] * J[1][1];
            const float h11 = J[0][1] * J[0][1] + J[1][1] * J[1][1];
            const float g0 = J[0][0] * e0 + J[1][0] * e1;
            const float g1 = J[0][1] * e0 + J[1][1] * e1;
            const float det = h00 * h11 - h01 * h01;
            if (!(det != 0.f)) break;
            sx = sx + (g0 * h11 - h01 * g1) / det;
            sy = sy + (h00 * g1 - h01 * g0) / det;
        }
        px[i] = sx * cam[0] + cam[2];
        py[i] = sy * cam[1] + cam[3];
    }
}
