"""Keypoint undistortion on the GPU: bit-exact with the CPU restatement for the
reference's two Gauss-Newton steps and for more, ragged and empty inputs,
all-zero distortion untouched, argument errors."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CAM = (1200.0, 1180.0, 800.0, 600.0)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,iters,k", [(2000, 2, (-0.12, 0.03, 0.0, 0.0, 0.0)),
                                       (100003, 2, (0.08, -0.02, 0.001, -0.0015, 0.004)),
                                       (777, 5, (-0.25, 0.08, -0.002, 0.001, -0.01)),
                                       (1, 2, (-0.1, 0.0, 0.0, 0.0, 0.0)),
                                       (300, 0, (-0.1, 0.0, 0.0, 0.0, 0.0))])
def test_undistort_bit_exact(orc, dev, n, iters, k):
    from rapidsfm_amd import ops

    _say("undistort", n, iters, k)
    rng = np.random.default_rng(n)
    px = rng.uniform(0, 1600, n).astype(np.float32)
    py = rng.uniform(0, 1200, n).astype(np.float32)
    ex, ey = orc.undistort(px, py, np.array(CAM, np.float32), np.array(k, np.float32), iters)
    gx, gy = torch.from_numpy(px).to(dev), torch.from_numpy(py).to(dev)
    ops.undistort_in_place(gx, gy, CAM, k, iters)
    assert np.array_equal(gx.cpu().numpy().view(np.uint32), ex.view(np.uint32))
    assert np.array_equal(gy.cpu().numpy().view(np.uint32), ey.view(np.uint32))


def test_zero_distortion_and_empty(dev):
    from rapidsfm_amd import ops

    _say("undistort zero/empty")
    px = torch.linspace(0, 1600, 999, device=dev)
    py = torch.linspace(0, 1200, 999, device=dev)
    x0, y0 = px.clone(), py.clone()
    ops.undistort_in_place(px, py, CAM, (0, 0, 0, 0, 0))
    assert torch.equal(px, x0) and torch.equal(py, y0)
    e = torch.empty(0, device=dev)
    ops.undistort_in_place(e, e.clone(), CAM, (-0.1, 0, 0, 0, 0))


def test_argument_errors(dev):
    from rapidsfm_amd import ops

    _say("undistort errors")
    px = torch.zeros(10, device=dev)
    with pytest.raises(RuntimeError):
        ops.undistort_in_place(px, px.clone(), CAM, (-0.1, 0, 0, 0, 0), 17)
    with pytest.raises(RuntimeError):
        ops.undistort_in_place(px, px.clone(), (0.0, 1.0, 0.0, 0.0), (-0.1, 0, 0, 0, 0))
    with pytest.raises(ValueError):
        ops.undistort_in_place(px, torch.zeros(9, device=dev), CAM, (-0.1, 0, 0, 0, 0))


def test_undistort_pt_pairs_bit_exact(orc, dev):
    from rapidsfm_amd import ops

    _say("undistort pt pairs, 15 tasks")
    rng = np.random.default_rng(77)
    probs, refs = [], []
    for i in range(15):
        n = int(rng.integers(0, 3000))
        pts = np.column_stack([rng.uniform(0, 1600, n), rng.uniform(0, 1200, n), rng.uniform(0, 1600, n),
                               rng.uniform(0, 1200, n)]).astype(np.float32)
        c0 = (1200.0 + i, 1190.0, 800.0, 600.0)
        c1 = (1100.0, 1105.0 + i, 790.0, 610.0)
        k0 = (-0.1 - 0.01 * i, 0.02, 0.0, 0.001, 0.0) if i % 3 else (0.0,) * 5
        k1 = (0.05, -0.01, 0.0005, 0.0, 0.002) if i % 4 else (0.0,) * 5
        live = n if i % 2 else max(0, n - 100)
        ndev = torch.tensor([live], dtype=torch.int32, device=dev)
        g = torch.from_numpy(pts).to(dev)
        probs.append((g, c0, c1, k0, k1, ndev))
        e = pts.copy()
        if live:
            e[:live, 0], e[:live, 1] = orc.undistort(pts[:live, 0], pts[:live, 1], np.array(c0, np.float32),
                                                     np.array(k0, np.float32), 2)
            e[:live, 2], e[:live, 3] = orc.undistort(pts[:live, 2], pts[:live, 3], np.array(c1, np.float32),
                                                     np.array(k1, np.float32), 2)
        refs.append(e)
    ops.undistort_pt_pairs(probs, 2)
    for (g, *_), e in zip(probs, refs):
        assert np.array_equal(g.cpu().numpy().view(np.uint32), e.view(np.uint32))
