"""4-point homography hypotheses on the GPU: counts and models bit-exact with the
FP32 oracle for the same sample indices; best hypothesis, inlier mask, gate and
edge cases."""
import numpy as np
import pytest
import torch

from tests.scenes import plane_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("seed,shape,H,t", [(1, (700, 60, 240), 1024, 3.0), (2, (3000, 500, 2000), 700, 2.0),
                                            (3, (4, 0, 0), 8, 1.0), (4, (200, 0, 3000), 2048, 4.0)])
def test_h4_bit_exact(orc, dev, seed, shape, H, t):
    from rapidsfm_amd import ops

    _say("ransac h4", seed, shape, H, t)
    pts, _, _, _ = plane_scene(seed, *shape, noise=0.5)
    n = len(pts)
    rng = np.random.default_rng(seed)
    idx = np.stack([rng.choice(n, 4, replace=False) for _ in range(H)]).astype(np.uint32)
    rc, rm = orc.h4_score(pts, idx, t, nthreads=8)
    (o,) = ops.ransac_h4([(torch.from_numpy(pts).to(dev), torch.from_numpy(idx.view(np.int32)).to(dev))], t,
                         want_models=True)
    assert np.array_equal(o["counts"].cpu().numpy().view(np.uint32), rc)
    assert np.array_equal(o["models"].cpu().numpy().reshape(H, 9).view(np.uint32), rm.view(np.uint32))
    hb = int(np.argmax(rc))
    best = o["best"].cpu().numpy().view(np.uint32)
    assert best[0] == hb and best[1] == rc[hb]
    if rc[hb] > 0:
        mask = orc.h_inlier_f32(rm[hb], pts, t).astype(np.uint8)
        assert np.array_equal(o["mask"].cpu().numpy(), mask) and int(mask.sum()) == rc[hb]


def test_h4_gate_device_count_and_errors(orc, dev):
    from rapidsfm_amd import ops

    _say("ransac h4 gate / device count")
    pts, _, _, _ = plane_scene(9, 800, 100, 300, noise=0.5)
    rng = np.random.default_rng(9)
    idx = np.stack([rng.choice(600, 4, replace=False) for _ in range(300)]).astype(np.uint32)
    idx[7, 1] = 5000  # out of range: rejected
    gate = (120.0 ** 2, 120.0 ** 2)
    live = 600
    rc, _ = orc.h4_score(pts[:live], idx, 3.0, gate[0], gate[1])
    pt = torch.from_numpy(pts).to(dev)
    nd = torch.tensor([live], dtype=torch.int32, device=dev)
    (o,) = ops.ransac_h4([(pt, torch.from_numpy(idx.view(np.int32)).to(dev), nd)], 3.0, min_sep_sq=gate)
    assert np.array_equal(o["counts"].cpu().numpy().view(np.uint32), rc)
    assert rc[7] == 0 and (rc == 0).sum() > 5
    assert not o["mask"].cpu().numpy()[live:].any()
    with pytest.raises(RuntimeError):
        ops.ransac_h4([(pt, torch.from_numpy(idx.view(np.int32)).to(dev))], -1.0)
