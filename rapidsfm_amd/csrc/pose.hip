// pose.hip - relative pose from a fundamental matrix or a homography: the
// (R, t) decompositions and the cheirality count that disambiguates them, one
// workgroup per image pair.
//
// Reference path: decomposeEpipolarity / decomposeEssentialMatrix
// (legacy/geometry.hpp:268-346: E = K1^T F K0, SVD, R = U W^T V^T or U W V^T,
// t = +-U.col(2), candidates in the order (R0,t0) (R0,-t0) (R1,t0) (R1,-t0)),
// isInFrontOfCam (ImagePair.cpp:760-766) over triangulate
// (legacy/geometry.cpp:146-206) and disambiguateSolutions
// (ImagePair.cpp:768-786: keep candidates with at least round(n * ratio) and
// at least 2 points in front of both cameras, most first).
// Homography branch: decomposeHomography (legacy/geometry.hpp:348-447, the
// SVD method of Ma, Soatto, Kosecka and Sastry: H = K1^-1 G K0 scaled by its
// middle singular value, V of H^T H with det V = +1, u1,2 = a v1 +- b v3,
// R = W U^T, N = v2 x u, T/d = (H - R) N, the four (R, N, T/d) of H and the
// four of -H, kept where N . T/d < 1), t = T/d normalised
// (ImagePair.cpp:810-816), then the same disambiguation.
//
// Own numerics: E and its SVD in FP64 on one lane (V from a cyclic Jacobi on
// E^T E, columns ordered by decreasing eigenvalue, u_k = E v_k / |E v_k| for
// the two largest, u3 = u1 x u2 and v3 = v1 x v2 so det U = det V = +1, which
// is the sign convention the reference enforces; for H the eigenvalues of
// H^T H divided by the middle one stand for the singular values of the
// normalised matrix); R, t and N rounded to FP32;
// the triangulation and depth tests in FP32 over the inliers, 256 lanes, with
// the 3x3 solve by the adjugate (A^-1 = adj(A) / det A). The CPU oracle
// repeats the same operations.
#include "internal.hpp"
#include "hgeom.hpp"
#include <algorithm>

namespace rsfm_amd {
namespace pose {

constexpr int kThreads = 256;
constexpr int kMaxTasks = 32;

typedef ByValue<DevTask, kMaxTasks> Batch;

__device__ inline uint32_t task_n(const DevTask& T) { return T.nDev ? min(*T.nDev, T.nHost) : T.nHost; }

// the four candidates, R row-major then t, FP32; false when E is degenerate
__device__ inline bool decompose(const float Fin[9], const float c0[4], const float c1[4], float out[4][12])
{
    double F[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) F[r][c] = (double)Fin[r * 3 + c];
    const double fx0 = c0[0], fy0 = c0[1], cx0 = c0[2], cy0 = c0[3];
    const double fx1 = c1[0], fy1 = c1[1], cx1 = c1[2], cy1 = c1[3];
    double G[3][3], E[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        G[a][0] = F[a][0] * fx0;
        G[a][1] = F[a][1] * fy0;
        G[a][2] = (F[a][0] * cx0 + F[a][1] * cy0) + F[a][2];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        E[0][c] = fx1 * G[0][c];
        E[1][c] = fy1 * G[1][c];
        E[2][c] = (cx1 * G[0][c] + cy1 * G[1][c]) + G[2][c];
    }
    double S[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] = (E[0][i] * E[0][j] + E[1][i] * E[1][j]) + E[2][i] * E[2][j];
    hgeom::jacobi3d(S, V);
    // columns by decreasing eigenvalue (first index on ties)
    int i0 = 0;
    if (S[1][1] > S[i0][i0]) i0 = 1;
    if (S[2][2] > S[i0][i0]) i0 = 2;
    int i1 = i0 == 0 ? 1 : 0;
    for (int k = 0; k < 3; k++)
        if (k != i0 && k != i1 && S[k][k] > S[i1][i1]) i1 = k;
    const double v1[3] = {V[0][i0], V[1][i0], V[2][i0]};
    const double v2[3] = {V[0][i1], V[1][i1], V[2][i1]};
    double u1[3], u2[3], u3[3], v3[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        u1[r] = (E[r][0] * v1[0] + E[r][1] * v1[1]) + E[r][2] * v1[2];
        u2[r] = (E[r][0] * v2[0] + E[r][1] * v2[1]) + E[r][2] * v2[2];
    }
    const double n1 = sqrt(hgeom::dot3(u1, u1)), n2 = sqrt(hgeom::dot3(u2, u2));
    if (!(n1 > 0.0) || !(n2 > 0.0) || !isfinite(n1) || !isfinite(n2)) return false;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        u1[r] = u1[r] / n1;
        u2[r] = u2[r] / n2;
    }
    hgeom::cross3(u1, u2, u3);
    hgeom::cross3(v1, v2, v3);
    // R0 = U W^T V^T = -u2 v1^T + u1 v2^T + u3 v3^T, R1 = U W V^T = u2 v1^T - u1 v2^T + u3 v3^T
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double a = u1[r] * v2[c] - u2[r] * v1[c];
            const double b = u3[r] * v3[c];
            const float R0 = (float)(a + b), R1 = (float)(b - a);
            out[0][r * 3 + c] = R0;
            out[1][r * 3 + c] = R0;
            out[2][r * 3 + c] = R1;
            out[3][r * 3 + c] = R1;
        }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float t = (float)u3[r];
        out[0][9 + r] = t;
        out[1][9 + r] = -t;
        out[2][9 + r] = t;
        out[3][9 + r] = -t;
    }
    return true;
}

// the valid homography candidates (N . T/d < 1) in the reference's order,
// compacted: R row-major then t = T/d / |T/d|, and N; returns their number
__device__ inline int decompose_h(const float Hin[9], const float c0[4], const float c1[4], float out[8][12],
                                  float nrm[8][3])
{
    double G[3][3], Hn[3][3];
    const double fx0 = c0[0], fy0 = c0[1], cx0 = c0[2], cy0 = c0[3];
    const double fx1 = c1[0], fy1 = c1[1], cx1 = c1[2], cy1 = c1[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double h0 = Hin[a * 3 + 0], h1 = Hin[a * 3 + 1], h2 = Hin[a * 3 + 2];
        G[a][0] = h0 * fx0;
        G[a][1] = h1 * fy0;
        G[a][2] = (h0 * cx0 + h1 * cy0) + h2;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        Hn[0][c] = (G[0][c] - cx1 * G[2][c]) / fx1;
        Hn[1][c] = (G[1][c] - cy1 * G[2][c]) / fy1;
        Hn[2][c] = G[2][c];
    }
    double S[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] = (Hn[0][i] * Hn[0][j] + Hn[1][i] * Hn[1][j]) + Hn[2][i] * Hn[2][j];
    hgeom::jacobi3d(S, V);
    int i0 = 0;
    if (S[1][1] > S[i0][i0]) i0 = 1;
    if (S[2][2] > S[i0][i0]) i0 = 2;
    int i1 = i0 == 0 ? 1 : 0;
    for (int k = 0; k < 3; k++)
        if (k != i0 && k != i1 && S[k][k] > S[i1][i1]) i1 = k;
    const int i2 = 3 - i0 - i1;
    const double lam1 = S[i1][i1];
    if (!(lam1 > 0.0) || !isfinite(S[i0][i0])) return 0;
    const double sc = sqrt(lam1);
    const double s0 = S[i0][i0] / lam1, s2 = S[i2][i2] / lam1;
    double v1[3] = {V[0][i0], V[1][i0], V[2][i0]};
    double v2[3] = {V[0][i1], V[1][i1], V[2][i1]};
    double v3[3] = {V[0][i2], V[1][i2], V[2][i2]};
    double x[3];
    hgeom::cross3(v2, v3, x);
    if (hgeom::dot3(v1, x) < 0.0) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            v1[r] = -v1[r];
            v2[r] = -v2[r];
            v3[r] = -v3[r];
        }
    }
    const double dd = s0 - s2;
    if (!(dd > 0.0)) return 0;
    const double cc = 1.0 / sqrt(dd);
    const double a = sqrt(1.0 - s2) * cc, b = sqrt(s0 - 1.0) * cc;
    double u[2][3], nv[2][3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        u[0][r] = a * v1[r] + b * v3[r];
        u[1][r] = a * v1[r] - b * v3[r];
    }
    hgeom::cross3(v2, u[0], nv[0]);
    hgeom::cross3(v2, u[1], nv[1]);
    int nOut = 0;
    for (int sgn = 0; sgn < 2; sgn++) {
        double Hs[3][3];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) Hs[r][c] = sgn ? -(Hn[r][c] / sc) : Hn[r][c] / sc;
        double w0[3];
#pragma unroll
        for (int r = 0; r < 3; r++) w0[r] = (Hs[r][0] * v2[0] + Hs[r][1] * v2[1]) + Hs[r][2] * v2[2];
        double R[2][3][3], T[2][3];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            double w1[3], w2[3];
#pragma unroll
            for (int r = 0; r < 3; r++) w1[r] = (Hs[r][0] * u[k][0] + Hs[r][1] * u[k][1]) + Hs[r][2] * u[k][2];
            hgeom::cross3(w0, w1, w2);
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) R[k][r][c] = (w0[r] * v2[c] + w1[r] * u[k][c]) + w2[r] * nv[k][c];
#pragma unroll
            for (int r = 0; r < 3; r++)
                T[k][r] = ((Hs[r][0] - R[k][r][0]) * nv[k][0] + (Hs[r][1] - R[k][r][1]) * nv[k][1]) +
                          (Hs[r][2] - R[k][r][2]) * nv[k][2];
        }
        // (R1, N1, T1), (R2, N2, T2), (R1, -N1, -T1), (R2, -N2, -T2); -N . -T = N . T
        for (int q = 0; q < 4; q++) {
            const int k = q & 1;
            if (!(hgeom::dot3(nv[k], T[k]) < 1.0)) continue;
            const double tn = sqrt(hgeom::dot3(T[k], T[k]));
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) out[nOut][r * 3 + c] = (float)R[k][r][c];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const double t = tn > 0.0 ? T[k][r] / tn : T[k][r];
                out[nOut][9 + r] = (float)(q < 2 ? t : -t);
                nrm[nOut][r] = (float)(q < 2 ? nv[k][r] : -nv[k][r]);
            }
            nOut++;
        }
    }
    return nOut;
}

// isInFrontOfCam for one correspondence under pose P (R row-major, t):
// triangulate (legacy/geometry.cpp:158-203) then depth > 0 in both cameras
__device__ inline bool in_front(const float P[12], float4 p, const float inv0[4], const float inv1[4])
{
    const float r0[3] = {inv0[0] * p.x + inv0[2], inv0[1] * p.y + inv0[3], 1.f};
    const float q1[3] = {inv1[0] * p.z + inv1[2], inv1[1] * p.w + inv1[3], 1.f};
    // ray1 = R^T q1, C1 = -R^T t
    float r1[3], C1[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r1[k] = (P[0 * 3 + k] * q1[0] + P[1 * 3 + k] * q1[1]) + P[2 * 3 + k] * q1[2];
        C1[k] = -((P[0 * 3 + k] * P[9] + P[1 * 3 + k] * P[10]) + P[2 * 3 + k] * P[11]);
    }
    const float l0 = sqrtf((r0[0] * r0[0] + r0[1] * r0[1]) + r0[2] * r0[2]);
    const float l1 = sqrtf((r1[0] * r1[0] + r1[1] * r1[1]) + r1[2] * r1[2]);
    float a[3], b[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a[k] = r0[k] / l0;
        b[k] = r1[k] / l1;
    }
    float A[3][3], M1[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float d = i == j ? 1.f : 0.f;
            M1[i][j] = b[i] * b[j] - d;
            A[i][j] = (a[i] * a[j] - d) + M1[i][j];
        }
    float rhs[3];
#pragma unroll
    for (int i = 0; i < 3; i++) rhs[i] = (M1[i][0] * C1[0] + M1[i][1] * C1[1]) + M1[i][2] * C1[2];
    // adjugate solve
    float adj[3][3];
    adj[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    adj[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    adj[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    adj[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    adj[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    adj[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    adj[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    adj[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    adj[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const float det = (A[0][0] * adj[0][0] + A[0][1] * adj[1][0]) + A[0][2] * adj[2][0];
    const float inv = 1.f / det;
    float X[3];
#pragma unroll
    for (int i = 0; i < 3; i++) X[i] = ((adj[i][0] * rhs[0] + adj[i][1] * rhs[1]) + adj[i][2] * rhs[2]) * inv;
    const float z1 = ((P[6] * X[0] + P[7] * X[1]) + P[8] * X[2]) + P[11];
    return X[2] > 0.f && z1 > 0.f;
}

// kKind 0: F (4 candidates, result [10]); 1: H (up to 8, result [19])
template <int kKind, class Src>
__global__ __launch_bounds__(kThreads) void pose_kernel(const Src batch, float ratio)
{
    constexpr int kC = kKind ? 8 : 4;
    const DevTask& T = batch[blockIdx.x];
    const uint32_t n = task_n(T);
    __shared__ float sP[kC][12], sN[kC][3];
    __shared__ int sNC;
    __shared__ uint32_t sCnt[kC], sIn;
    if (threadIdx.x == 0) {
        float M[9];
        for (int q = 0; q < 9; q++) M[q] = T.M[q];
        float P[kC][12], Nv[kC][3];
        int nc;
        if constexpr (kKind == 0) {
            nc = decompose(M, T.cam0, T.cam1, P) ? 4 : 0;
        } else {
            nc = decompose_h(M, T.cam0, T.cam1, P, Nv);
        }
        for (int k = 0; k < kC; k++) {
            for (int q = 0; q < 12; q++) sP[k][q] = k < nc ? P[k][q] : 0.f;
            if constexpr (kKind == 1)
                for (int q = 0; q < 3; q++) sN[k][q] = k < nc ? Nv[k][q] : 0.f;
            sCnt[k] = 0;
        }
        sNC = nc;
        sIn = 0;
    }
    __syncthreads();
    // inverse cameras: f' = 1 / f, c' = -c f' (PinHoleCameraTemplate::inverse, Types.h:236-241)
    float inv0[4], inv1[4];
    inv0[0] = 1.f / T.cam0[0];
    inv0[1] = 1.f / T.cam0[1];
    inv0[2] = -T.cam0[2] * inv0[0];
    inv0[3] = -T.cam0[3] * inv0[1];
    inv1[0] = 1.f / T.cam1[0];
    inv1[1] = 1.f / T.cam1[1];
    inv1[2] = -T.cam1[2] * inv1[0];
    inv1[3] = -T.cam1[3] * inv1[1];
    const int nc = sNC;
    uint32_t c[kC], in = 0;
#pragma unroll
    for (int k = 0; k < kC; k++) c[k] = 0u;
    if (nc > 0) {
        float P[kC][12];
#pragma unroll
        for (int k = 0; k < kC; k++)
#pragma unroll
            for (int q = 0; q < 12; q++) P[k][q] = sP[k][q];
        for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
            if (T.mask && !T.mask[i]) continue;
            const float4 p = T.pts[i];
            in++;
#pragma unroll
            for (int k = 0; k < kC; k++)
                if (k < nc) c[k] += in_front(P[k], p, inv0, inv1) ? 1u : 0u;
        }
    }
    for (int k = 0; k < kC; k++)
        if (c[k]) atomicAdd(&sCnt[k], c[k]);
    if (in) atomicAdd(&sIn, in);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t order[kC];
        for (int k = 0; k < kC; k++) order[k] = k;
        // most points in front first, lower candidate index on ties; only the nc valid ones
        for (int a = 1; a < nc; a++)
            for (int b = a; b > 0 && sCnt[order[b]] > sCnt[order[b - 1]]; b--) {
                const uint32_t t = order[b];
                order[b] = order[b - 1];
                order[b - 1] = t;
            }
        const uint32_t minCnt = (uint32_t)roundf((float)sIn * ratio);
        uint32_t kept = 0;
        while ((int)kept < nc && sCnt[order[kept]] >= minCnt && sCnt[order[kept]] >= 2) kept++;
        for (int k = 0; k < kC; k++) {
            T.result[k] = sCnt[k];
            T.result[kC + k] = order[k];
        }
        T.result[2 * kC] = kept;
        T.result[2 * kC + 1] = sIn;
        if constexpr (kKind == 1) T.result[2 * kC + 2] = (uint32_t)nc;
    }
    if (threadIdx.x < kC * 12) T.poses[threadIdx.x] = sP[threadIdx.x / 12][threadIdx.x % 12];
    if constexpr (kKind == 1)
        if (T.normals && threadIdx.x < kC * 3) T.normals[threadIdx.x] = sN[threadIdx.x / 3][threadIdx.x % 3];
}

}  // namespace pose
}  // namespace rsfm_amd

using namespace rsfm_amd;

namespace rsfm_amd {
namespace pose {

template <int kKind, class Task>
int make_task_t(const Task& t, DevTask& d)
{
    d = DevTask{};
    if constexpr (kKind == 0) {
        d.M = t.F;
    } else {
        d.M = t.H;
        d.normals = t.normals;
    }
    RSFM_RETURN_IF(d.M == nullptr || t.poses == nullptr || t.result == nullptr, RSFM_ERR_INVALID,
                   "pose: the model, poses and result are required");
    RSFM_RETURN_IF(t.nbPts > 0 && t.pts == nullptr, RSFM_ERR_INVALID, "pose: pts is null");
    RSFM_RETURN_IF(reinterpret_cast<uintptr_t>(t.pts) % 16, RSFM_ERR_INVALID, "pose: pts must be 16-byte aligned");
    RSFM_RETURN_IF(!(t.cam0.fx != 0.f) || !(t.cam0.fy != 0.f) || !(t.cam1.fx != 0.f) || !(t.cam1.fy != 0.f),
                   RSFM_ERR_INVALID, "pose: focal lengths must be nonzero");
    const float c0[4] = {t.cam0.fx, t.cam0.fy, t.cam0.cx, t.cam0.cy};
    const float c1[4] = {t.cam1.fx, t.cam1.fy, t.cam1.cx, t.cam1.cy};
    for (int k = 0; k < 4; k++) {
        d.cam0[k] = c0[k];
        d.cam1[k] = c1[k];
    }
    d.pts = reinterpret_cast<const float4*>(t.pts);
    d.nDev = t.nbPtsDev;
    d.nHost = t.nbPts;
    d.mask = t.inlierMask;
    d.poses = t.poses;
    d.result = t.result;
    return RSFM_OK;
}

int make_task(const rsfm_pose_task& t, DevTask& d) { return make_task_t<0>(t, d); }
int make_task(const rsfm_pose_h_task& t, DevTask& d) { return make_task_t<1>(t, d); }

int check_ratio(float frontRatio)
{
    RSFM_RETURN_IF(!(frontRatio >= 0.f) || frontRatio > 1.f, RSFM_ERR_INVALID, "pose: frontRatio must be in [0, 1]");
    return RSFM_OK;
}

int launch_pose_table(const DevTask* tasks, uint32_t nTasks, int kind, float frontRatio, hipStream_t stream)
{
    RSFM_CHECK(check_ratio(frontRatio));
    if (nTasks == 0) return RSFM_OK;
    const InTable<DevTask> src{tasks};
    if (kind == 0) pose_kernel<0><<<nTasks, kThreads, 0, stream>>>(src, frontRatio);
    else pose_kernel<1><<<nTasks, kThreads, 0, stream>>>(src, frontRatio);
    return check_launch("pose table");
}

}  // namespace pose
}  // namespace rsfm_amd

namespace {
template <int kKind, class Task>
int launch_pose(const Task* tasks, size_t nbTasks, float frontRatio, hipStream_t stream, const char* what)
{
    using namespace pose;
    RSFM_RETURN_IF(nbTasks > 0 && tasks == nullptr, RSFM_ERR_INVALID, "pose: tasks is null");
    RSFM_CHECK(check_ratio(frontRatio));
    for (size_t b = 0; b < nbTasks; b += kMaxTasks) {
        const size_t n = std::min<size_t>(kMaxTasks, nbTasks - b);
        Batch batch{};
        for (size_t i = 0; i < n; i++) RSFM_CHECK(make_task(tasks[b + i], batch.t[i]));
        pose_kernel<kKind><<<(uint32_t)n, kThreads, 0, stream>>>(batch, frontRatio);
        RSFM_CHECK(check_launch(what));
    }
    return RSFM_OK;
}
}  // namespace

extern "C" int rsfm_pose_from_f(const rsfm_pose_task* tasks, size_t nbTasks, float frontRatio, hipStream_t stream)
{
    return launch_pose<0>(tasks, nbTasks, frontRatio, stream, "pose from F");
}

extern "C" int rsfm_pose_from_h(const rsfm_pose_h_task* tasks, size_t nbTasks, float frontRatio, hipStream_t stream)
{
    return launch_pose<1>(tasks, nbTasks, frontRatio, stream, "pose from H");
}
