"""DEGENSAC finisher on the GPU: F, H, counts, source, best sample and both
inlier masks bit-exact with the CPU restatement for the same sample indices
and counts; chained after the GPU 7-point scorer; batches, device counts and
edge cases."""
import numpy as np
import pytest
import torch

from tests.scenes import plane_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _samples(n, H, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.choice(n, 7, replace=False) for _ in range(H)]).astype(np.uint32)


def _check(o, ref, n):
    assert np.array_equal(o["result"].cpu().numpy().view(np.uint32), ref["result"])
    assert np.array_equal(o["F"].cpu().numpy().reshape(9).view(np.uint32), ref["F"].view(np.uint32))
    assert np.array_equal(o["H"].cpu().numpy().reshape(9).view(np.uint32), ref["H"].view(np.uint32))
    assert np.array_equal(o["mask_f"].cpu().numpy()[:n], ref["mask_f"])
    assert np.array_equal(o["mask_h"].cpu().numpy()[:n], ref["mask_h"])


@pytest.mark.parametrize("seed,mode,shape,K,S", [
    (0, 0, (700, 60, 240), 8, 256),
    (1, 1, (700, 60, 240), 8, 256),
    (2, 0, (3000, 400, 1600), 16, 300),
    (3, 1, (0, 900, 400), 4, 64),
    (4, 0, (20000, 2000, 8000), 4, 512),
])
def test_refine_bit_exact(orc, dev, seed, mode, shape, K, S):
    from rapidsfm_amd import ops

    _say("degensac", seed, mode, shape, K, S)
    t_f = 1.5 if mode == 0 else 1.0
    pts, lab, _, _ = plane_scene(seed, *shape, noise=0.3)
    n = pts.shape[0]
    idx = _samples(n, 512, seed)
    c, _ = orc.f7_score(pts, idx, t_f, mode, nthreads=8)
    ref = orc.degensac_refine(pts, idx, c, t_f, 3.0, mode, K, S, 99 + seed)
    (o,) = ops.degensac_refine([(torch.from_numpy(pts).to(dev), torch.from_numpy(idx.view(np.int32)).to(dev),
                                 torch.from_numpy(c.view(np.int32)).to(dev))], t_f, 3.0, mode=mode, nb_candidates=K,
                               nb_plane_samples=S, seed=99 + seed)
    _check(o, ref, n)
    if shape[0] > 0:
        assert ref["result"][1] > 0.9 * shape[0]


def test_chained_after_gpu_scorer(orc, dev):
    from rapidsfm_amd import ops

    _say("degensac chained")
    pts, lab, _, _ = plane_scene(7, 1500, 120, 500, noise=0.3)
    n = pts.shape[0]
    idx = _samples(n, 1024, 7)
    pt = torch.from_numpy(pts).to(dev)
    it = torch.from_numpy(idx.view(np.int32)).to(dev)
    (s,) = ops.ransac_f7([(pt, it)], 1.5, mode=0)
    (o,) = ops.degensac_refine([(pt, it, s["counts"])], 1.5, 3.0, mode=0, nb_candidates=8, nb_plane_samples=256,
                               seed=5)
    c, _ = orc.f7_score(pts, idx, 1.5, 0, nthreads=8)
    assert np.array_equal(s["counts"].cpu().numpy().view(np.uint32), c)
    _check(o, orc.degensac_refine(pts, idx, c, 1.5, 3.0, 0, 8, 256, 5), n)


def test_batch_device_count_and_edges(orc, dev):
    from rapidsfm_amd import ops

    _say("degensac batch")
    probs, refs, ns = [], [], []
    for k, shape in enumerate([(600, 50, 200), (0, 500, 300), (800, 80, 100)]):
        pts, _, _, _ = plane_scene(30 + k, *shape, noise=0.3)
        n = pts.shape[0]
        idx = _samples(n, 256, k)
        c, _ = orc.f7_score(pts, idx, 1.5, 0, nthreads=8)
        probs.append((torch.from_numpy(pts).to(dev), torch.from_numpy(idx.view(np.int32)).to(dev),
                      torch.from_numpy(c.view(np.int32)).to(dev)))
        refs.append(orc.degensac_refine(pts, idx, c, 1.5, 3.0, 0, 6, 128, 11))
        ns.append(n)
    # a device count below the capacity: only the first 700 points are live
    pts, _, _, _ = plane_scene(40, 600, 60, 240, noise=0.3)
    live = 700
    idx = _samples(live, 256, 4)
    c, _ = orc.f7_score(pts[:live], idx, 1.5, 0, nthreads=8)
    probs.append((torch.from_numpy(pts).to(dev), torch.from_numpy(idx.view(np.int32)).to(dev),
                  torch.from_numpy(c.view(np.int32)).to(dev), torch.tensor([live], dtype=torch.int32, device=dev)))
    r = orc.degensac_refine(pts[:live], idx, c, 1.5, 3.0, 0, 6, 128, 11)
    r["mask_f"] = np.concatenate([r["mask_f"], np.zeros(pts.shape[0] - live, np.uint8)])
    r["mask_h"] = np.concatenate([r["mask_h"], np.zeros(pts.shape[0] - live, np.uint8)])
    refs.append(r)
    ns.append(pts.shape[0])
    # all counts zero, and fewer than 7 points
    probs.append((probs[0][0], probs[0][1], torch.zeros_like(probs[0][2])))
    refs.append(None)
    ns.append(None)
    few = torch.from_numpy(pts[:6].copy()).to(dev)
    probs.append((few, torch.from_numpy((idx[:3] % 6).view(np.int32)).to(dev),
                  torch.ones(3, dtype=torch.int32, device=dev)))
    refs.append(None)
    ns.append(None)
    outs = ops.degensac_refine(probs, 1.5, 3.0, mode=0, nb_candidates=6, nb_plane_samples=128, seed=11)
    for o, ref, n in zip(outs, refs, ns):
        if ref is None:
            assert list(o["result"].cpu().numpy().view(np.uint32)) == [0, 0, 0, 0xFFFFFFFF]
            assert not o["F"].any() and not o["mask_f"].any() and not o["mask_h"].any()
        else:
            _check(o, ref, n)


def test_argument_errors(dev):
    from rapidsfm_amd import ops

    pts = torch.zeros((100, 4), dtype=torch.float32, device=dev)
    idx = torch.zeros((4, 7), dtype=torch.int32, device=dev)
    cnt = torch.zeros(4, dtype=torch.int32, device=dev)
    with pytest.raises(RuntimeError):
        ops.degensac_refine([(pts, idx, cnt)], 1.0, 3.0, nb_candidates=0)
    with pytest.raises(RuntimeError):
        ops.degensac_refine([(pts, idx, cnt)], 1.0, 3.0, nb_plane_samples=5000)
    with pytest.raises(RuntimeError):
        ops.degensac_refine([(pts, idx, cnt)], -1.0, 3.0)
    big = torch.zeros((70000, 4), dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError):
        ops.degensac_refine([(big, idx, cnt)], 1.0, 3.0)
