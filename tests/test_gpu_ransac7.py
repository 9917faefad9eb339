"""7-point hypotheses on the GPU: per-hypothesis counts and models bit-exact with the
FP32 oracle for the same sample indices; best hypothesis and inlier mask."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("mode,t,seed,n,n_out,H", [(0, 4.0, 1, 2000, 800, 1024), (1, 1.0, 2, 4096, 1638, 768),
                                                   (1, 1.0, 3, 300, 0, 256), (0, 4.0, 4, 7, 0, 16)])
def test_f7_bit_exact(orc, dev, mode, t, seed, n, n_out, H):
    from rapidsfm_amd import ops

    _say("ransac f7", mode, t, seed, n, n_out, H)
    pts, _, _ = orc.gen_two_view(seed, n, n_out)
    rng = np.random.default_rng(seed)
    idx = np.stack([rng.choice(n, 7, replace=False) for _ in range(H)]).astype(np.uint32)
    ref_c, ref_m = orc.f7_score(pts, idx, t, mode)
    pt = torch.from_numpy(pts).to(dev)
    it = torch.from_numpy(idx.view(np.int32)).to(dev)
    (o,) = ops.ransac_f7([(pt, it)], t, mode=mode, want_models=True)
    c = o["counts"].cpu().numpy().view(np.uint32)
    m = o["models"].cpu().numpy().reshape(H, 9)
    assert np.array_equal(c, ref_c)
    assert np.array_equal(m.view(np.uint32), ref_m.view(np.uint32))
    best = o["best"].cpu().numpy().view(np.uint32)
    hb = int(np.argmax(ref_c))  # first maximum = lowest index
    assert best[0] == hb and best[1] == ref_c[hb]
    if ref_c[hb] > 0:
        F = ref_m[hb]
        mask = orc.f_inlier_f32(F, pts, t, mode).astype(np.uint8)
        assert np.array_equal(o["mask"].cpu().numpy(), mask)
        assert int(mask.sum()) == ref_c[hb]
    # without stored models the best sample is refitted: same best and mask
    (o2,) = ops.ransac_f7([(pt, it)], t, mode=mode, want_models=False)
    assert np.array_equal(o2["best"].cpu().numpy(), o["best"].cpu().numpy())
    assert np.array_equal(o2["mask"].cpu().numpy(), o["mask"].cpu().numpy())


def test_f7_gate_and_edge_cases(orc, dev):
    from rapidsfm_amd import ops

    _say("ransac f7 edges")
    pts, _, _ = orc.gen_two_view(9, 1000, 200)
    rng = np.random.default_rng(0)
    idx = np.stack([rng.choice(1000, 7, replace=False) for _ in range(300)]).astype(np.uint32)
    idx[5, 2] = 5000  # out of range: rejected
    gate = (150.0 ** 2, 150.0 ** 2)
    ref_c, _ = orc.f7_score(pts, idx, 2.0, 0, gate[0], gate[1])
    (o,) = ops.ransac_f7([(torch.from_numpy(pts).to(dev), torch.from_numpy(idx.view(np.int32)).to(dev))], 2.0,
                         mode=0, min_sep_sq=gate)
    assert np.array_equal(o["counts"].cpu().numpy().view(np.uint32), ref_c)
    assert ref_c[5] == 0 and (ref_c == 0).sum() > 5
    # fewer than 7 points: nothing is fitted
    few = torch.from_numpy(pts[:6]).to(dev)
    (o2,) = ops.ransac_f7([(few, torch.from_numpy(idx[:4].view(np.int32) % 6).to(dev))], 2.0, mode=0)
    assert not o2["counts"].any() and int(o2["best"][1]) == 0
