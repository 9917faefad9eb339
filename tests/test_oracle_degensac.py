"""Pins of the DEGENSAC restatement (oracle/degensac.c) by the mathematics.

Refits reproduce noise-free F and H and agree with an independent SVD solution
(numpy) of the same normalised least-squares problems on noisy data; H from F
and F from H satisfy their defining relations; the homography test matches a
direct FP64 evaluation of checkProjectiveTransformation's inequality away from
its boundary; on dominant-plane scenes the finisher recovers the off-plane
inliers that the plain 7-point best misses.
"""
import numpy as np
import pytest

from tests.scenes import plane_scene, same_up_to_scale


def _hartley(p):
    c = p.mean(axis=0)
    d = np.sqrt(((p - c) ** 2).sum(axis=1)).mean()
    s = np.sqrt(2.0) / d
    T = np.array([[s, 0, -s * c[0]], [0, s, -s * c[1]], [0, 0, 1.0]])
    return (p - c) * s, T


def _svd_f(pts):
    """normalised 8-point least squares by SVD, rank 2 by SVD (independent route)"""
    p = pts.astype(np.float64)
    a, T0 = _hartley(p[:, :2])
    b, T1 = _hartley(p[:, 2:])
    A = np.column_stack([b[:, 0] * a[:, 0], b[:, 0] * a[:, 1], b[:, 0], b[:, 1] * a[:, 0], b[:, 1] * a[:, 1], b[:, 1],
                         a[:, 0], a[:, 1], np.ones(len(p))])  # x1^T F x0 with F row-major
    F = np.linalg.svd(A)[2][-1].reshape(3, 3)
    U, S, Vt = np.linalg.svd(F)
    F = U @ np.diag([S[0], S[1], 0.0]) @ Vt
    return T1.T @ F @ T0


def _svd_h(pts):
    p = pts.astype(np.float64)
    a, T0 = _hartley(p[:, :2])
    b, T1 = _hartley(p[:, 2:])
    rows = []
    for (x, y), (u, v) in zip(a, b):
        rows.append([-x, -y, -1, 0, 0, 0, x * u, y * u, u])
        rows.append([0, 0, 0, -x, -y, -1, x * v, y * v, v])
    Hn = np.linalg.svd(np.array(rows))[2][-1].reshape(3, 3)
    return np.linalg.inv(T1) @ Hn @ T0


def _skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def test_refits_noise_free(orc):
    pts, lab, F, H = plane_scene(1, 200, 100, 0)
    off = np.nonzero(lab == 1)[0]
    pl = np.nonzero(lab == 0)[0]
    assert same_up_to_scale(orc.fit_ls_f(pts, off[:50]), F) < 1e-6
    assert same_up_to_scale(orc.fit_ls_f(pts, np.arange(128)), F) < 1e-6
    assert same_up_to_scale(orc.fit_ls_h(pts, pl[:40]), H) < 1e-6
    assert same_up_to_scale(orc.fit_ls_h(pts, pl[:4]), H) < 1e-5
    # too few points: no fit
    assert orc.fit_ls_f(pts, off[:7]) is None and orc.fit_ls_h(pts, pl[:3]) is None


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_refits_match_svd(orc, seed):
    pts, lab, F, H = plane_scene(seed, 300, 300, 0, noise=0.5)
    gen = np.nonzero(lab >= 0)[0][:128]
    Ff = orc.fit_ls_f(pts, gen)
    assert same_up_to_scale(Ff, _svd_f(pts[gen])) < 2e-6
    pl = np.nonzero(lab == 0)[0][:100]
    assert same_up_to_scale(orc.fit_ls_h(pts, pl), _svd_h(pts[pl])) < 2e-6
    # rank 2
    s = np.linalg.svd(Ff.astype(np.float64), compute_uv=False)
    assert s[2] < 1e-6 * s[0]


@pytest.mark.parametrize("seed", [6, 7])
def test_h_from_f(orc, seed):
    pts, lab, F, H = plane_scene(seed, 100, 50, 0)
    pl = np.nonzero(lab == 0)[0]
    H3 = orc.h_from_f(F.astype(np.float32), pts[pl[:3]])
    assert same_up_to_scale(H3, H) < 1e-5  # the plane through three plane points
    # three off-plane points span another plane: H maps them exactly and F ~ [e']x H
    tri = pts[np.nonzero(lab == 1)[0][:3]]
    Hx = orc.h_from_f(F.astype(np.float32), tri).astype(np.float64)
    for q in tri.astype(np.float64):
        y = Hx @ np.array([q[0], q[1], 1.0])
        assert np.abs(y[:2] / y[2] - q[2:]).max() < 1e-2
    e = np.linalg.svd(F.T)[2][-1]
    assert same_up_to_scale(_skew(e) @ Hx, F) < 1e-5
    # collinear triplet in image 0: no homography
    col = tri.copy()
    col[:, :2] = [[100.0, 200.0], [300.0, 400.0], [500.0, 600.0]]
    assert orc.h_from_f(F.astype(np.float32), col) is None


@pytest.mark.parametrize("seed", [8, 9])
def test_f_from_h(orc, seed):
    pts, lab, F, H = plane_scene(seed, 100, 50, 0)
    off = np.nonzero(lab == 1)[0]
    assert same_up_to_scale(orc.f_from_h(H.astype(np.float32), pts[off[0]], pts[off[1]]), F) < 1e-5
    ptsn, labn, _, Hn = plane_scene(seed, 100, 50, 0, noise=1.0)
    pa, pb = ptsn[np.nonzero(labn == 1)[0][:2]]
    Hf = Hn.astype(np.float32)
    Fh = orc.f_from_h(Hf, pa, pb).astype(np.float64)
    for q in (pa, pb):
        x0 = np.array([q[0], q[1], 1.0])
        x1 = np.array([q[2], q[3], 1.0])
        assert abs(x1 @ Fh @ x0) < 1e-6 * np.linalg.norm(x1) * np.linalg.norm(x0)
    s = np.linalg.svd(Fh, compute_uv=False)
    assert s[2] < 1e-6 * s[0]
    # H^T [e']x H is antisymmetric; elementwise bound from the magnitudes summed
    M = Hf.astype(np.float64).T @ Fh
    bound = np.abs(Hf.astype(np.float64)).T @ np.abs(Fh)
    assert (np.abs(M + M.T) <= 1e-6 * (bound + bound.T)).all()


def test_h_inlier_formula(orc):
    rng = np.random.default_rng(11)
    pts, lab, F, H = plane_scene(10, 300, 300, 200, noise=1.0)
    Hf = (H * rng.choice([-1.0, 1.0]) * 3.0).astype(np.float32)
    for t in (0.5, 2.0, 8.0):
        got = orc.h_inlier_f32(Hf, pts, t)
        h = Hf.astype(np.float64).reshape(9)
        p = pts.astype(np.float64)
        z = h[8] + h[6] * p[:, 0] + h[7] * p[:, 1]
        sc = np.minimum(z * z, abs(h[0] * h[4] - h[3] * h[1]))
        lhs = (h[2] + h[0] * p[:, 0] + h[1] * p[:, 1] - p[:, 2] * z) ** 2 + (
            h[5] + h[3] * p[:, 0] + h[4] * p[:, 1] - p[:, 3] * z) ** 2
        rhs = t * t * sc
        clear = np.abs(lhs - rhs) > 1e-3 * rhs
        assert clear.sum() > 0.9 * len(p)
        assert np.array_equal(got[clear], (lhs < rhs)[clear])
        assert 0 < got.sum() < len(p)


@pytest.mark.parametrize("seed,mode", [(0, 0), (1, 1), (2, 0), (3, 1)])
def test_finisher_recovers_off_plane(orc, seed, mode):
    t_f = 1.5 if mode == 0 else 1.0
    pts, lab, F, H = plane_scene(seed, 700, 60, 240, noise=0.3)
    n = pts.shape[0]
    rng = np.random.default_rng(seed)
    idx = np.stack([rng.choice(n, 7, replace=False) for _ in range(512)]).astype(np.uint32)
    c, _ = orc.f7_score(pts, idx, t_f, mode, nthreads=4)
    r = orc.degensac_refine(pts, idx, c, t_f, 3.0, mode, 8, 256, 77)
    res = r["result"]
    assert res[0] >= c.max()  # the first candidate's optimisation never loses inliers
    mf = r["mask_f"].astype(bool)
    assert int(mf.sum()) == res[0]
    assert (mf & (lab == 1)).sum() >= 0.95 * (lab == 1).sum()
    assert (mf & (lab == 2)).sum() <= 5
    assert res[1] >= 0.95 * (lab == 0).sum()  # the dominant plane's homography
    mh = r["mask_h"].astype(bool)
    assert int(mh.sum()) == res[1] and (mh & (lab == 0)).sum() >= 0.95 * (lab == 0).sum()
    assert same_up_to_scale(r["F"], F) < 5e-3


def test_finisher_general_scene_and_edges(orc):
    # no dominant plane: the optimised 7-point F wins, no plane found
    pts, lab, F, H = plane_scene(21, 0, 600, 300, noise=0.3)
    n = pts.shape[0]
    rng = np.random.default_rng(1)
    idx = np.stack([rng.choice(n, 7, replace=False) for _ in range(512)]).astype(np.uint32)
    c, _ = orc.f7_score(pts, idx, 1.5, 0, nthreads=4)
    r = orc.degensac_refine(pts, idx, c, 1.5, 3.0, 0, 4, 64, 5)
    assert r["result"][2] == 0 and r["result"][0] >= 0.97 * 600
    assert c[r["result"][3]] > 0
    # all counts zero / too few points: nothing
    z = orc.degensac_refine(pts, idx, np.zeros_like(c), 1.5, 3.0, 0, 4, 64, 5)
    assert list(z["result"]) == [0, 0, 0, 0xFFFFFFFF] and not z["F"].any() and not z["mask_f"].any()
    few = orc.degensac_refine(pts[:6], idx[:3] % 6, np.ones(3, np.uint32), 1.5, 3.0, 0, 4, 64, 5)
    assert list(few["result"]) == [0, 0, 0, 0xFFFFFFFF]
