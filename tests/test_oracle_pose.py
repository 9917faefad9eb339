"""Pins of the relative-pose restatement (oracle/pose.c) by geometry.

Noise-free scenes: one candidate is the true (R, t/|t|) and has every point in
front of both cameras, the rotations are orthonormal with det +1, [t]x R is
proportional to K1^T F K0, triangulation returns the true 3D points; with noise
and outliers the true candidate still leads and is the only one kept.
"""
import numpy as np
import pytest

from tests.scenes import posed_scene, same_up_to_scale


def _skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def _K(c):
    return np.array([[c[0], 0, c[2]], [0, c[1], c[3]], [0, 0, 1.0]])


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_decomposition_contains_true_pose(orc, seed):
    pts, c0, c1, R, t, X, F = posed_scene(seed, 200, focal1=1100.0 + 50 * seed)
    P = orc.pose_decompose(F.astype(np.float32), c0, c1)
    E = _K(c1).T @ F @ _K(c0)
    hits = 0
    for k in range(4):
        Rk, tk = P[k, :9].reshape(3, 3).astype(np.float64), P[k, 9:].astype(np.float64)
        assert np.abs(Rk @ Rk.T - np.eye(3)).max() < 1e-5 and abs(np.linalg.det(Rk) - 1) < 1e-5
        assert abs(np.linalg.norm(tk) - 1) < 1e-6
        assert same_up_to_scale(_skew(tk) @ Rk, E) < 1e-5
        hits += np.abs(Rk - R).max() < 1e-5 and np.abs(tk - t).max() < 1e-5
    assert hits == 1


@pytest.mark.parametrize("seed", [5, 6])
def test_triangulation_and_cheirality(orc, seed):
    pts, c0, c1, R, t, X, F = posed_scene(seed, 150)
    poses, res = orc.pose_from_f(F.astype(np.float32), c0, c1, pts)
    best = res[4]
    assert res[best] == 150 and res[8] == 1 and res[9] == 150
    assert sorted(res[:4])[-2] < 150
    for i in range(0, 150, 17):
        Xi, front = orc.pose_triangulate(poses[best], pts[i], c0, c1)
        assert front and np.abs(Xi - X[i]).max() < 1e-3 * np.abs(X[i]).max()  # FP32, depth up to 20 baselines


@pytest.mark.parametrize("seed", [7, 8])
def test_noise_outliers_and_mask(orc, seed):
    pts, c0, c1, R, t, X, F = posed_scene(seed, 400, 200, noise=0.5)
    mask = np.zeros(600, np.uint8)
    mask[:400] = 1
    poses, res = orc.pose_from_f(F.astype(np.float32), c0, c1, pts, mask)
    best = res[4]
    Rb, tb = poses[best, :9].reshape(3, 3), poses[best, 9:]
    assert np.abs(Rb - R).max() < 1e-5 and np.abs(tb - t).max() < 1e-5
    assert res[9] == 400 and res[best] >= 0.95 * 400 and res[8] == 1
    # a ratio of 0 keeps every candidate with at least two points in front
    _, res0 = orc.pose_from_f(F.astype(np.float32), c0, c1, pts, mask, 0.0)
    assert res0[8] == sum(int(res0[k] >= 2) for k in range(4))
