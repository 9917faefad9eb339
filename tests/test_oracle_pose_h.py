"""Pins of the homography-branch pose restatement (oracle or_pose_decompose_h /
or_pose_from_h).

Noise-free plane scenes: every valid candidate is a decomposition of the
camera-normalised homography, H ~ R + (T/d) N^T up to the sign and scale of H,
with R orthonormal, det R = +1, |N| = 1 and |t| = 1, and satisfies the
reference's filter N . T/d < 1; one candidate is the true (R, t, plane normal);
it alone has every point in front of both cameras and is kept first. With noise,
outliers outside the mask and either sign of H the true pose still leads.
"""
import numpy as np
import pytest

from tests.scenes import posed_plane_scene


def _norm_h(H, c0, c1):
    K0 = np.array([[c0[0], 0, c0[2]], [0, c0[1], c0[3]], [0, 0, 1.0]])
    K1 = np.array([[c1[0], 0, c1[2]], [0, c1[1], c1[3]], [0, 0, 1.0]])
    Hn = np.linalg.inv(K1) @ H.astype(np.float64) @ K0
    return Hn / np.linalg.svd(Hn, compute_uv=False)[1]


def _true_index(P, N, R, t, nrm):
    err = [max(np.abs(p[:9].reshape(3, 3) - R).max(), np.abs(p[9:] - t).max(), np.abs(nv - nrm).max())
           for p, nv in zip(P, N)]
    return int(np.argmin(err)), min(err)


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_candidates_decompose_h_and_include_truth(orc, seed):
    pts, c0, c1, R, t, nrm, d, H = posed_plane_scene(seed, 200, focal1=1100.0 + 40 * seed)
    Hf = H.astype(np.float32)
    P, N = orc.pose_decompose_h(Hf, c0, c1)
    assert 2 <= len(P) <= 8
    Hn = _norm_h(Hf, c0, c1)
    for p, nv in zip(P, N):
        Rm = p[:9].reshape(3, 3).astype(np.float64)
        assert np.abs(Rm @ Rm.T - np.eye(3)).max() < 1e-5 and abs(np.linalg.det(Rm) - 1) < 1e-5
        assert abs(np.linalg.norm(p[9:]) - 1) < 1e-5 and abs(np.linalg.norm(nv) - 1) < 1e-5
        # T/d = (H - R) N is parallel to t: H = R + s t N^T for the sign of H that the candidate used
        ok = False
        for sg in (1.0, -1.0):
            D = sg * Hn - Rm
            s = float(p[9:].astype(np.float64) @ D @ nv.astype(np.float64))
            if np.abs(D - s * np.outer(p[9:], nv)).max() < 1e-4:
                ok = ok or s * float(nv @ p[9:]) < 1.0 + 1e-6
        assert ok
    k, err = _true_index(P, N, R, t, nrm)
    assert err < 1e-4


@pytest.mark.parametrize("seed,sign", [(6, 1.0), (7, -1.0), (8, 1.0)])
def test_cheirality_keeps_the_true_pose(orc, seed, sign):
    pts, c0, c1, R, t, nrm, d, H = posed_plane_scene(seed, 300)
    Hf = (sign * H).astype(np.float32)
    P, N, res = orc.pose_from_h(Hf, c0, c1, pts, None, 0.95)
    nc = res[18]
    k, err = _true_index(P[:nc], N[:nc], R, t, nrm)
    assert err < 1e-4
    assert res[k] == 300 and res[17] == 300
    assert res[8] == k and res[16] == 1
    # every other candidate loses points (the mirrored solution puts some behind a camera)
    assert all(res[j] < 300 for j in range(nc) if j != k)
    # unused slots are zero and ordered after the valid ones
    assert not P[nc:].any() and list(res[8 + nc:16]) == list(range(nc, 8))


def test_noise_outliers_and_mask(orc):
    pts, c0, c1, R, t, nrm, d, H = posed_plane_scene(9, 800, n_out=400, noise=0.7)
    mask = np.zeros(len(pts), np.uint8)
    mask[:800] = 1
    P, N, res = orc.pose_from_h(H.astype(np.float32), c0, c1, pts, mask, 0.9)
    assert res[17] == 800
    best = res[8]
    assert np.abs(P[best, :9].reshape(3, 3) - R).max() < 1e-3 and np.abs(P[best, 9:] - t).max() < 1e-3
    assert res[16] >= 1


def test_pure_rotation_gives_no_candidates(orc):
    # H = K R K^-1: all singular values of the normalised homography are 1
    c = np.array([1000.0, 1000.0, 640.0, 480.0], np.float32)
    K = np.array([[1000.0, 0, 640.0], [0, 1000.0, 480.0], [0, 0, 1]])
    H = (K @ np.eye(3) @ np.linalg.inv(K)).astype(np.float32)
    P, N = orc.pose_decompose_h(H, c, c)
    assert len(P) == 0
