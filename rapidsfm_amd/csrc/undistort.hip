// undistort.hip - undistortion of keypoint coordinates before pair solving.
//
// Reference: undistortInPlace / undistort (distortion.cpp:63-272, the model of
// DistortParams: k1, k2, p1, p2, k3, value = x (1 + k1 r^2 + k2 r^4 + k3 r^6)
// plus the tangential terms, its Jacobian, Gauss-Newton steps solving
// J^T J delta = J^T (target - value)). The reference's in-place driver drops
// the result of undistort (distortion.cpp:238), so its output is the input
// mapped to normalised coordinates and back; this kernel applies the
// undistortion it computes. FP32, one thread per point, the expressions in
// the reference's order without contraction, the 2x2 solve by Cramer's rule.
// Elementwise and HBM-bound: 8 bytes read and 8 written per point per array
// pair.
#include "internal.hpp"
#include <algorithm>

namespace rsfm_amd {
namespace undist {

__device__ inline void value_jacobian(const Params& d, float x0, float x1, float& v0, float& v1, float J[2][2])
{
    const float r2 = x0 * x0 + x1 * x1;
    const float scale = ((1.f + d.k1 * r2) + d.k2 * (r2 * r2)) + d.k3 * (r2 * r2 * r2);
    v0 = x0 * scale + ((2.f * d.p1 * (x0 * x1) + d.p2 * r2) + 2.f * d.p2 * (x0 * x0));
    v1 = x1 * scale + ((2.f * d.p2 * (x0 * x1) + d.p1 * r2) + 2.f * d.p1 * (x1 * x1));
    const float sj = (d.k1 + 2.f * d.k2 * r2) + 3.f * d.k3 * (r2 * r2);
    const float j01 = ((2.f * x0 * x1 * sj) + 2.f * d.p1 * x0) + 2.f * d.p2 * x1;
    J[0][0] = ((scale + 2.f * sj * (x0 * x0)) + 2.f * d.p1 * x1) + 6.f * d.p2 * x0;
    J[0][1] = j01;
    J[1][0] = j01;
    J[1][1] = ((scale + 2.f * sj * (x1 * x1)) + 6.f * d.p1 * x1) + 2.f * d.p2 * x0;
}

// one pixel (x, y) through inverse pinhole, Gauss-Newton, pinhole
__device__ inline void undistort_point(const Params& d, float& x, float& y)
{
    const float tx = x * d.ifx + d.icx, ty = y * d.ify + d.icy;
    float sx = tx, sy = ty;
    for (uint32_t it = 0; it < d.iters; it++) {
        float v0, v1, J[2][2];
        value_jacobian(d, sx, sy, v0, v1, J);
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
    x = sx * d.fx + d.cx;
    y = sy * d.fy + d.cy;
}

__global__ void undistort_kernel(float* __restrict__ px, float* __restrict__ py, uint32_t n, const Params d)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = px[i], y = py[i];
    undistort_point(d, x, y);
    px[i] = x;
    py[i] = y;
}

// keypoint pairs (x0, y0, x1, y1): the first point with camera 0, the second
// with camera 1 (an identity Params skips a side); one thread per pair,
// blockIdx.y = task
template <class Src>
__global__ void undistort_pairs_kernel(const Src tasks)
{
    const PairTask& T = tasks[blockIdx.y];
    const uint32_t n = T.nDev ? min(*T.nDev, T.nHost) : T.nHost;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = T.pts[i];
        if (T.on0) undistort_point(T.d0, p.x, p.y);
        if (T.on1) undistort_point(T.d1, p.z, p.w);
        T.pts[i] = p;
    }
}

}  // namespace undist
}  // namespace rsfm_amd

using namespace rsfm_amd;

namespace rsfm_amd {
namespace undist {

int make_params(const rsfm_real_camera& camera, uint32_t nbIters, Params& d, bool& active)
{
    RSFM_RETURN_IF(nbIters > 16, RSFM_ERR_INVALID, "undistort: nbIters must be <= 16");
    const rsfm_camera& c = camera.pinHole;
    RSFM_RETURN_IF(!(c.fx != 0.f) || !(c.fy != 0.f), RSFM_ERR_INVALID, "undistort: focal lengths must be nonzero");
    const float* k = camera.distortion;
    active = !(k[0] == 0.f && k[1] == 0.f && k[2] == 0.f && k[3] == 0.f && k[4] == 0.f);
    d.fx = c.fx;
    d.fy = c.fy;
    d.cx = c.cx;
    d.cy = c.cy;
    d.ifx = 1.f / c.fx;
    d.ify = 1.f / c.fy;
    d.icx = -c.cx * d.ifx;
    d.icy = -c.cy * d.ify;
    d.k1 = k[0];
    d.k2 = k[1];
    d.p1 = k[2];
    d.p2 = k[3];
    d.k3 = k[4];
    d.iters = nbIters;
    return RSFM_OK;
}

int launch_pairs_table(const PairTask* tasks, uint32_t nTasks, uint32_t maxN, hipStream_t stream)
{
    if (nTasks == 0 || maxN == 0) return RSFM_OK;
    undistort_pairs_kernel<<<dim3(std::min<uint32_t>(div_up(maxN, 256), 64), nTasks), 256, 0, stream>>>(
        InTable<PairTask>{tasks});
    return check_launch("undistort pairs");
}

}  // namespace undist
}  // namespace rsfm_amd

extern "C" int rsfm_undistort_in_place(float* px, float* py, size_t nbPts, const rsfm_real_camera* camera,
                                       uint32_t nbIters, hipStream_t stream)
{
    RSFM_RETURN_IF(camera == nullptr, RSFM_ERR_INVALID, "undistort: camera is null");
    RSFM_RETURN_IF(nbPts > 0 && (px == nullptr || py == nullptr), RSFM_ERR_INVALID, "undistort: px/py null");
    RSFM_RETURN_IF(nbPts > 0xFFFFFFFFull, RSFM_ERR_UNSUPPORTED, "undistort: at most 2^32 - 1 points per call");
    undist::Params d;
    bool active = false;
    RSFM_CHECK(undist::make_params(*camera, nbIters, d, active));
    if (nbPts == 0 || !active) return RSFM_OK;
    undist::undistort_kernel<<<div_up((uint32_t)nbPts, 256), 256, 0, stream>>>(px, py, (uint32_t)nbPts, d);
    return check_launch("undistort");
}

extern "C" int rsfm_undistort_pt_pairs(const rsfm_undistort_task* tasks, size_t nbTasks, uint32_t nbIters,
                                       hipStream_t stream)
{
    using namespace undist;
    constexpr int kMaxTasks = 12;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "undistort pairs: tasks is null");
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        ByValue<PairTask, kMaxTasks> batch{};
        uint32_t maxN = 0;
        for (size_t i = 0; i < n; i++) {
            const rsfm_undistort_task& t = tasks[b + i];
            RSFM_RETURN_IF(t.nbPts > 0 && t.pts == nullptr, RSFM_ERR_INVALID, "undistort pairs: pts is null");
            RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16, RSFM_ERR_INVALID,
                           "undistort pairs: pts must be 16-byte aligned");
            PairTask& d = batch.t[i];
            bool a0 = false, a1 = false;
            RSFM_CHECK(make_params(t.cam0, nbIters, d.d0, a0));
            RSFM_CHECK(make_params(t.cam1, nbIters, d.d1, a1));
            d.pts = reinterpret_cast<float4*>(t.pts);
            d.nDev = t.nbPtsDev;
            d.nHost = (a0 || a1) ? t.nbPts : 0u;
            d.on0 = a0;
            d.on1 = a1;
            maxN = std::max(maxN, d.nHost);
        }
        if (maxN == 0) continue;
        undistort_pairs_kernel<<<dim3(std::min<uint32_t>(div_up(maxN, 256), 64), (uint32_t)n), 256, 0, stream>>>(batch);
        RSFM_CHECK(check_launch("undistort pairs"));
    }
    return RSFM_OK;
}
