"""Relative pose from F on the GPU: the four candidates, the front counts, their
order and the kept number bit-exact with the CPU restatement; batches, masks,
device counts and argument errors."""
import numpy as np
import pytest
import torch

from tests.scenes import posed_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _check(o, ref_poses, ref_res):
    assert np.array_equal(o["poses"].cpu().numpy().view(np.uint32), ref_poses.view(np.uint32))
    assert np.array_equal(o["result"].cpu().numpy().view(np.uint32), ref_res)


@pytest.mark.parametrize("seed,n,n_out,noise,ratio", [(1, 300, 0, 0.0, 0.95), (2, 2000, 800, 0.5, 0.95),
                                                      (3, 5000, 3000, 1.0, 0.5), (4, 1, 0, 0.0, 0.95)])
def test_pose_bit_exact(orc, dev, seed, n, n_out, noise, ratio):
    from rapidsfm_amd import ops

    _say("pose", seed, n, n_out, noise, ratio)
    pts, c0, c1, R, t, X, F = posed_scene(seed, n, n_out, noise=noise, focal1=1150.0)
    mask = np.zeros(len(pts), np.uint8)
    mask[:n] = 1
    Ff = F.astype(np.float32)
    rp, rr = orc.pose_from_f(Ff, c0, c1, pts, mask, ratio)
    (o,) = ops.pose_from_f([(torch.from_numpy(Ff).to(dev), c0, c1, torch.from_numpy(pts).to(dev),
                             torch.from_numpy(mask).to(dev))], ratio)
    _check(o, rp, rr)
    if n >= 100:
        best = rr[4]
        assert np.abs(rp[best, :9].reshape(3, 3) - R).max() < 1e-4 and rr[8] >= 1


def test_pose_batch_device_count_no_mask(orc, dev):
    from rapidsfm_amd import ops

    _say("pose batch")
    probs, refs = [], []
    for k in range(5):
        pts, c0, c1, R, t, X, F = posed_scene(20 + k, 500 + 300 * k, 100, noise=0.3)
        Ff = F.astype(np.float32)
        live = len(pts) - 50
        refs.append(orc.pose_from_f(Ff, c0, c1, pts[:live], None, 0.95))
        probs.append((torch.from_numpy(Ff).to(dev), c0, c1, torch.from_numpy(pts).to(dev), None,
                      torch.tensor([live], dtype=torch.int32, device=dev)))
    outs = ops.pose_from_f(probs, 0.95)
    for o, (rp, rr) in zip(outs, refs):
        _check(o, rp, rr)
    # a zero F: nothing decomposes, nothing is kept
    (z,) = ops.pose_from_f([(torch.zeros((3, 3), device=dev), c0, c1, probs[0][3])], 0.95)
    r = z["result"].cpu().numpy()
    assert r[8] == 0 and not z["poses"].any()
    with pytest.raises(RuntimeError):
        ops.pose_from_f([(probs[0][0], c0, c1, probs[0][3])], 1.5)
