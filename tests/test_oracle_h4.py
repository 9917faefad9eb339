"""Pins of the 4-point homography restatement (oracle or_h4_fit_f32 / or_h4_score).

The FP32 fit reproduces the plane's true H from four noise-free plane points and
agrees with an independent FP64 solve of the same four-point DLT (numpy SVD,
Hartley normalised) on noisy points; it maps its four points onto their partners;
coincident samples are rejected; the scorer
counts exactly the correspondences the pinned homography test accepts.
"""
import numpy as np
import pytest

from tests.scenes import plane_scene, same_up_to_scale


def _dlt4_f64(p):
    p = p.astype(np.float64)
    rows = []
    for x, y, u, v in p:
        rows.append([-x, -y, -1, 0, 0, 0, x * u, y * u, u])
        rows.append([0, 0, 0, -x, -y, -1, x * v, y * v, v])
    return np.linalg.svd(np.array(rows))[2][-1].reshape(3, 3)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fit_recovers_plane_homography(orc, seed):
    pts, lab, F, H = plane_scene(seed, 60, 0, 0)
    H4 = orc.h4_fit_f32(pts[:4])
    assert same_up_to_scale(H4, H) < 2e-4


@pytest.mark.parametrize("seed", [4, 5, 6])
def test_fit_matches_fp64_dlt_and_maps_its_points(orc, seed):
    pts, lab, F, H = plane_scene(seed, 60, 0, 0, noise=2.0)
    p4 = pts[:4]
    H4 = orc.h4_fit_f32(p4)
    assert same_up_to_scale(H4, _dlt4_f64(p4)) < 2e-4
    for x, y, u, v in p4.astype(np.float64):
        q = H4.astype(np.float64) @ np.array([x, y, 1.0])
        assert np.abs(q[:2] / q[2] - (u, v)).max() < 0.05


def test_degenerate_samples_rejected(orc):
    pts, _, _, _ = plane_scene(7, 60, 0, 0)
    dup = pts[:4].copy()
    dup[:] = dup[0]
    assert orc.h4_fit_f32(dup) is None


def test_scorer_counts_the_pinned_test(orc):
    pts, lab, F, H = plane_scene(8, 400, 100, 300, noise=0.5)
    rng = np.random.default_rng(8)
    idx = np.stack([rng.choice(len(pts), 4, replace=False) for _ in range(200)]).astype(np.uint32)
    c, m = orc.h4_score(pts, idx, 3.0, nthreads=4)
    for h in range(0, 200, 37):
        Hm = orc.h4_fit_f32(pts[idx[h]])
        if Hm is None:
            assert c[h] == 0
            continue
        assert np.array_equal(m[h].view(np.uint32), Hm.reshape(9).view(np.uint32))
        assert c[h] == orc.h_inlier_f32(Hm, pts, 3.0).sum()
    # a sample of four plane points explains the plane
    pl = np.nonzero(lab == 0)[0]
    corners = [(0, 0), (1600, 0), (0, 1200), (1600, 1200)]  # the plane point nearest each image corner
    plane = np.array([pl[np.argmin(((pts[pl, :2] - c) ** 2).sum(1))] for c in corners], np.uint32)
    cp, _ = orc.h4_score(pts, plane[None, :], 3.0)
    assert cp[0] >= 0.95 * (lab == 0).sum()
    # the sanity gate rejects samples whose points are closer than the gate
    cg, _ = orc.h4_score(pts, idx, 3.0, 1e12, 0.0)
    assert not cg.any()
