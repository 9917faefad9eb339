"""Pins of the keypoint undistortion restatement (oracle or_undistort_in_place).

The forward model is checked against the Brown-Conrady formula written out in
FP64 (radial k1 r^2 + k2 r^4 + k3 r^6, tangential p1, p2 in OpenCV's order);
two Gauss-Newton steps agree with the same two steps run in FP64 with a
central-difference Jacobian (a wrong Jacobian term changes the iterates
although not the fixed point); enough steps reach the FP64 fixed point, so that
distorting the result gives the input back; all-zero distortion leaves the
points untouched bit for bit.
"""
import numpy as np
import pytest

CAM = np.array([1200.0, 1180.0, 800.0, 600.0], np.float32)


def _distort_f64(k, x, y):
    k1, k2, p1, p2, k3 = [float(v) for v in k]
    r2 = x * x + y * y
    radial = 1 + k1 * r2 + k2 * r2 ** 2 + k3 * r2 ** 3
    return (x * radial + 2 * p1 * x * y + p2 * (r2 + 2 * x * x),
            y * radial + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y)


def _gn_f64(k, tx, ty, iters):
    sx, sy = tx.copy(), ty.copy()
    h = 1e-6
    for _ in range(iters):
        v0, v1 = _distort_f64(k, sx, sy)
        a0, a1 = _distort_f64(k, sx + h, sy)
        b0, b1 = _distort_f64(k, sx - h, sy)
        c0, c1 = _distort_f64(k, sx, sy + h)
        d0, d1 = _distort_f64(k, sx, sy - h)
        j00, j10 = (a0 - b0) / (2 * h), (a1 - b1) / (2 * h)
        j01, j11 = (c0 - d0) / (2 * h), (c1 - d1) / (2 * h)
        e0, e1 = tx - v0, ty - v1
        det = j00 * j11 - j01 * j10
        sx = sx + (j11 * e0 - j01 * e1) / det
        sy = sy + (j00 * e1 - j10 * e0) / det
    return sx, sy


def _points(seed, n=2000):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0, 1600, n).astype(np.float32), rng.uniform(0, 1200, n).astype(np.float32))


DISTORTIONS = [
    (-0.12, 0.03, 0.0, 0.0, 0.0),
    (0.08, -0.02, 0.001, -0.0015, 0.004),
    (-0.25, 0.08, -0.002, 0.001, -0.01),
]


@pytest.mark.parametrize("k", DISTORTIONS)
def test_forward_model_matches_brown_conrady(orc, k):
    rng = np.random.default_rng(1)
    for x, y in rng.uniform(-0.7, 0.7, (50, 2)).astype(np.float32):
        got = orc.distort_f32(np.array(k, np.float32), float(x), float(y))
        want = _distort_f64(np.array(k, np.float32).astype(np.float64), np.float64(x), np.float64(y))
        assert np.abs(got - np.array(want)).max() < 2e-6


@pytest.mark.parametrize("k", DISTORTIONS)
def test_two_steps_match_fp64_gauss_newton(orc, k):
    px, py = _points(2)
    kk = np.array(k, np.float32)
    ux, uy = orc.undistort(px, py, CAM, kk, iters=2)
    c = CAM.astype(np.float64)
    tx = (px.astype(np.float64) - c[2]) / c[0]
    ty = (py.astype(np.float64) - c[3]) / c[1]
    # FP64 normal equations J^T J d = J^T e with a square J are J d = e
    sx, sy = _gn_f64(kk.astype(np.float64), tx, ty, 2)
    assert np.abs(ux - (sx * c[0] + c[2])).max() < 2e-3
    assert np.abs(uy - (sy * c[1] + c[3])).max() < 2e-3


@pytest.mark.parametrize("k", DISTORTIONS)
def test_converged_result_distorts_back_to_the_input(orc, k):
    px, py = _points(3)
    kk = np.array(k, np.float32)
    ux, uy = orc.undistort(px, py, CAM, kk, iters=8)
    c = CAM.astype(np.float64)
    sx = (ux.astype(np.float64) - c[2]) / c[0]
    sy = (uy.astype(np.float64) - c[3]) / c[1]
    dx, dy = _distort_f64(kk.astype(np.float64), sx, sy)
    assert np.abs(dx * c[0] + c[2] - px).max() < 2e-3
    assert np.abs(dy * c[1] + c[3] - py).max() < 2e-3
    # and the correction is real: points move by up to tens of pixels
    assert np.abs(ux - px).max() > 5.0


def test_zero_distortion_is_untouched(orc):
    px, py = _points(4)
    ux, uy = orc.undistort(px, py, CAM, np.zeros(5, np.float32), iters=2)
    assert np.array_equal(ux.view(np.uint32), px.view(np.uint32))
    assert np.array_equal(uy.view(np.uint32), py.view(np.uint32))


def test_zero_iterations_is_the_pinhole_round_trip(orc):
    px, py = _points(5)
    ux, uy = orc.undistort(px, py, CAM, np.array(DISTORTIONS[0], np.float32), iters=0)
    assert np.abs(ux - px).max() < 1e-3 and np.abs(uy - py).max() < 1e-3
