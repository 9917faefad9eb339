"""Relative pose from H on the GPU: the valid candidates, their normals, the
front counts, order, kept number and candidate count bit-exact with the CPU
restatement; batches beyond one launch, masks, device counts, either sign of H,
a pure rotation (no candidates) and argument errors."""
import numpy as np
import pytest
import torch

from tests.scenes import posed_plane_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _check(o, rp, rn, rr):
    assert np.array_equal(o["result"].cpu().numpy().view(np.uint32), rr)
    nc = int(rr[18])
    assert np.array_equal(o["poses"].cpu().numpy()[:nc].view(np.uint32), rp[:nc].view(np.uint32))
    assert np.array_equal(o["normals"].cpu().numpy()[:nc].view(np.uint32), rn[:nc].view(np.uint32))


@pytest.mark.parametrize("seed,n,n_out,noise,sign,ratio", [(1, 300, 0, 0.0, 1.0, 0.95), (2, 2000, 800, 0.5, -1.0, 0.95),
                                                           (3, 5000, 3000, 1.0, 1.0, 0.5), (4, 1, 0, 0.0, 1.0, 0.95)])
def test_pose_h_bit_exact(orc, dev, seed, n, n_out, noise, sign, ratio):
    from rapidsfm_amd import ops

    _say("pose_h", seed, n, n_out, noise, sign, ratio)
    pts, c0, c1, R, t, nrm, d, H = posed_plane_scene(seed, n, n_out, noise=noise, focal1=1150.0)
    mask = np.zeros(len(pts), np.uint8)
    mask[:n] = 1
    Hf = (sign * H).astype(np.float32)
    rp, rn, rr = orc.pose_from_h(Hf, c0, c1, pts, mask, ratio)
    (o,) = ops.pose_from_h([(torch.from_numpy(Hf).to(dev), c0, c1, torch.from_numpy(pts).to(dev),
                             torch.from_numpy(mask).to(dev))], ratio)
    _check(o, rp, rn, rr)
    if n >= 100:
        best = rr[8]
        assert np.abs(rp[best, :9].reshape(3, 3) - R).max() < 1e-3 and rr[16] >= 1


def test_pose_h_batch_and_device_counts(orc, dev):
    from rapidsfm_amd import ops

    _say("pose_h batch 40")
    probs, refs = [], []
    for s in range(40):
        pts, c0, c1, R, t, nrm, d, H = posed_plane_scene(100 + s, 150 + 7 * s, 20, noise=0.3)
        ndev = torch.tensor([100 + s], dtype=torch.int32, device=dev)
        Hf = H.astype(np.float32)
        probs.append((torch.from_numpy(Hf).to(dev), c0, c1, torch.from_numpy(pts).to(dev), None, ndev))
        refs.append(orc.pose_from_h(Hf, c0, c1, pts[:100 + s], None, 0.95))
    outs = ops.pose_from_h(probs, 0.95)
    for o, (rp, rn, rr) in zip(outs, refs):
        _check(o, rp, rn, rr)


def test_pose_h_pure_rotation(orc, dev):
    from rapidsfm_amd import ops

    _say("pose_h pure rotation")
    c = np.array([1000.0, 1000.0, 640.0, 480.0], np.float32)
    H = np.eye(3, dtype=np.float32)
    pts = torch.tensor([[10.0, 20.0, 10.0, 20.0]] * 8, device=dev)
    (o,) = ops.pose_from_h([(torch.from_numpy(H).to(dev), c, c, pts)])
    rp, rn, rr = orc.pose_from_h(H, c, c, pts.cpu().numpy())
    _check(o, rp, rn, rr)
    assert rr[18] == 0 and rr[16] == 0


def test_pose_h_errors(dev):
    from rapidsfm_amd import ops

    _say("pose_h errors")
    pts = torch.zeros((4, 4), device=dev)
    H = torch.eye(3, device=dev)
    with pytest.raises(RuntimeError):
        ops.pose_from_h([(H, (0.0, 1.0, 0.0, 0.0), (1.0, 1.0, 0.0, 0.0), pts)])
    with pytest.raises(RuntimeError):
        ops.pose_from_h([(H, (1.0, 1.0, 0.0, 0.0), (1.0, 1.0, 0.0, 0.0), pts)], front_ratio=1.5)
